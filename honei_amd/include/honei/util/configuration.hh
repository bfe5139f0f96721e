// honei/util/configuration.hh -- runtime configuration keys (`honeirc`).
//
// Same lookup and key syntax as HONEI's Configuration
// (honei/util/configuration.cc:97-200): "key = value" lines, '#' comments,
// read from $HONEI_CONFIG or ./honeirc on first use.  Keys used by this
// build: ell::threads (lanes per row of SparseMatrixELL built from a
// SparseMatrix, default 1).
#pragma once

#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

namespace honei
{
    class Configuration
    {
        private:
            std::map<std::string, std::string> _values;
            std::mutex _mutex;

            Configuration()
            {
                const char * env = std::getenv("HONEI_CONFIG");
                std::ifstream f(env ? env : "honeirc");
                std::string line;
                while (std::getline(f, line))
                {
                    const size_t hash = line.find('#');
                    if (hash != std::string::npos)
                        line.erase(hash);
                    const size_t eq = line.find('=');
                    if (eq == std::string::npos)
                        continue;
                    std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
                    if (! k.empty())
                        _values[k] = v;
                }
            }

            static std::string trim(const std::string & s)
            {
                const size_t b = s.find_first_not_of(" \t\r\n");
                if (b == std::string::npos)
                    return "";
                const size_t e = s.find_last_not_of(" \t\r\n");
                return s.substr(b, e - b + 1);
            }

        public:
            static Configuration * instance()
            {
                static Configuration c;
                return &c;
            }

            int get_value(const std::string & name, int default_value)
            {
                std::lock_guard<std::mutex> g(_mutex);
                auto it = _values.find(name);
                return it == _values.end() ? default_value : std::atoi(it->second.c_str());
            }

            std::string get_value(const std::string & name, const std::string & default_value)
            {
                std::lock_guard<std::mutex> g(_mutex);
                auto it = _values.find(name);
                return it == _values.end() ? default_value : it->second;
            }

            void set_value(const std::string & name, int value)
            {
                std::lock_guard<std::mutex> g(_mutex);
                _values[name] = std::to_string(value);
            }

            void set_value(const std::string & name, const std::string & value)
            {
                std::lock_guard<std::mutex> g(_mutex);
                _values[name] = value;
            }
    };
}
