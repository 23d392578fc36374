// Tangent build of oracle/lag2d_oracle_q2.c (see dual.hpp).
#include "dual.hpp"
#define double Dual
#include "../lag2d_oracle_q2.c"
#undef double
