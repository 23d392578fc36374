#include "lag2d_host.h"
int upload_extra(lag2d_ctx &) { return 0; }
int launch_rhs_q2(lag2d_ctx *, lag2d::Dev &, const double *, double, int) { return LAG2D_EUNSUPPORTED; }
