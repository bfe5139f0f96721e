// honei/math/operator_list.hh -- an ordered list of operators that runs them
// in turn (HONEI's honei/math/operator_list.hh:30-88).  The list owns the
// operators; copies share them, as the reference's list of raw pointers
// copied by value does.
#pragma once

#include <honei/math/operator.hh>

#include <memory>
#include <string>
#include <vector>

namespace honei
{
    class OperatorList
    {
        private:
            std::shared_ptr<std::vector<std::unique_ptr<Operator>>> _ops;

        public:
            OperatorList() : _ops(std::make_shared<std::vector<std::unique_ptr<Operator>>>()) {}

            void push_back(Operator * op) { _ops->emplace_back(op); }
            unsigned long size() const { return _ops->size(); }
            Operator * operator[](unsigned long i) const { return (*_ops)[i].get(); }

            /// runs every operator in order
            void value()
            {
                for (auto & op : *_ops)
                    op->value();
            }

            std::string to_string() const
            {
                std::string r;
                for (auto & op : *_ops)
                    r += op->to_string() + "\n";
                return r;
            }
    };
}
