// honei/la/scale.hh -- include path kept alike to HONEI; the templates are in operations.hh.
#pragma once
#include <honei/la/operations.hh>
