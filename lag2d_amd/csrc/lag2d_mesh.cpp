// Mesh-derived tables built on the host at context creation.
#include <hip/hip_runtime.h>

#include <vector>

#include "lag2d_host.h"

// Straight cells: each node appears at exactly one corner of each of its
// cells, and the reference scatters corner contributions cell by cell in
// increasing cell id -- triangles (Getrhslagmctria.f90:1186-1199), then quads
// (Limiterlag.f90:1011-1016), as the cells are numbered -- which is the order
// of the elements-around-points lists.  Entry = cell*4 + local vertex.
int build_node_corners(lag2d_ctx &x, const std::vector<int> &ipq, const std::vector<int> &esuv1,
                       const std::vector<int> &esuv2, std::vector<int> &n2c, std::vector<int> &n2c_off)
{
    const int nc = x.nc, np = x.np, nv = x.nvq;
    n2c_off.assign(np + 1, 0);
    n2c.clear();
    for (int p = 0; p < np; ++p) {
        int prev = -1;
        for (int k = esuv2[p]; k < esuv2[p + 1]; ++k) {
            int ie = esuv1[k];
            if (ie < prev) return LAG2D_EINVAL;  // element lists must be ascending
            prev = ie;
            int iv = -1;
            const int nvc = ie < x.ntria ? (nv == 4 ? 3 : 6) : (nv == 4 ? 4 : 9);
            for (int j = 0; j < nvc; ++j)
                if (ipq[(size_t)j * nc + ie] == p) iv = j;
            if (iv < 0) return LAG2D_EINVAL;
            if (nv == 4) {
                n2c.push_back(ie * 4 + iv);
            } else if (ie < x.ntria) {
                // curved triangle: its sub-triangle corners at node p, in
                // the order getriem_triasubg3 scatters them
                // (Getrhslaggauss.f90:12540-12585), before every quad
                static const int iptsg[4][3] = {{0, 3, 5}, {3, 4, 5}, {3, 1, 4}, {5, 4, 2}};
                for (int isg = 0; isg < 4; ++isg)
                    for (int ivsg = 0; ivsg < 3; ++ivsg)
                        if (iptsg[isg][ivsg] == iv) n2c.push_back(ie * 16 + isg * 4 + ivsg);
            } else {
                // sub-cell corners of this cell at node p, in the order the
                // sub-cell solver scatters them (Getrhslaggauss.f90:10639-10858)
                static const int ipqsg[4][4] = {{0, 4, 8, 7}, {4, 1, 5, 8}, {8, 5, 2, 6}, {7, 8, 6, 3}};
                for (int isg = 0; isg < 4; ++isg)
                    for (int ivsg = 0; ivsg < 4; ++ivsg)
                        if (ipqsg[isg][ivsg] == iv) n2c.push_back(ie * 16 + isg * 4 + ivsg);
            }
        }
        n2c_off[p + 1] = (int)n2c.size();
    }
    return 0;
}
