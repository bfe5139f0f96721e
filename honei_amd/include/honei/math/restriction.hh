// honei/math/restriction.hh -- include path kept alike to HONEI; the templates are in solvers.hh.
#pragma once
#include <honei/math/solvers.hh>
