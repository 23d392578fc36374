// Counting build of oracle/lag2d_oracle_q2.c (see flop.hpp), with the
// counter accessors.
#include "flop.hpp"

unsigned long long g_ops[16];
int g_phase;
long long g_weno_cells;

#define double Flop
#include "../lag2d_oracle_q2.c"
#undef double

extern "C" void orc_opcount(unsigned long long *ops, long long *weno_cells, int reset)
{
    for (int k = 0; k < 16; ++k) ops[k] = g_ops[k];
    *weno_cells = g_weno_cells;
    if (reset) {
        for (int k = 0; k < 16; ++k) g_ops[k] = 0;
        g_weno_cells = 0;
    }
}
