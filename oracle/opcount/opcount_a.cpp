// Counting build of oracle/lag2d_oracle.c (see flop.hpp).
#include "flop.hpp"
#define double Flop
#include "../lag2d_oracle.c"
