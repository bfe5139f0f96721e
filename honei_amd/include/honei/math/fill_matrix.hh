// honei/math/fill_matrix.hh -- include path kept alike to HONEI; the code is in io.hh.
#pragma once
#include <honei/math/io.hh>
