// Host side of the C ABI (include/lag2d.h): context, uploads in the
// reference's array layouts, and the per-stage launch sequence of the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lag2d.h"
#include "lag2d_dev.h"
#include "lag2d_kernels.h"
#include "lag2d_host.h"

using namespace lag2d;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(LAG2D_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" const char *lag2d_last_error(void) { return g_err.c_str(); }

// For the other translation units of the library (lag2d_noalloc.hip).
int lag2d_set_error(int code, const std::string &msg) { return fail(code, msg); }

template <class T>
static int dalloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(LAG2D_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

static inline int nblocks(int n) { return (n + LAG2D_BLOCK - 1) / LAG2D_BLOCK; }
static void launch_cellvol(lag2d_ctx *x, double *xs, double *ys, int set_centre, bool timed);

lag2d_ctx::~lag2d_ctx()
{
    for (auto &g : graphs)
        if (g.exec) hipGraphExecDestroy(g.exec);
    for (void *p : allocs) hipFree(p);
    for (auto &t : timers) {
        for (auto &ev : t.ev) {
            hipEventDestroy(ev.first);
            hipEventDestroy(ev.second);
        }
    }
    if (stream) hipStreamDestroy(stream);
}

template <class T>
int lag2d_ctx::alloc(T **p, size_t n)
{
    int rc = dalloc(p, n);
    if (!rc) allocs.push_back((void *)*p);
    return rc;
}

// Runs of an ensemble (one run for lag2d_create).
struct RunSpec {
    int nrun = 1;
    std::vector<int> coff, noff;
    std::vector<double> tend, dtfix, tacru;
};

static int create_impl(const lag2d_config *cfg, int ntria, int nquad, int npoin, int nbfac, const int *iptri,
                       const int *ipqua, const int *bface, const int *esuv1, const int *esuv2, const RunSpec &rs,
                       lag2d_ctx **out)
{
    if (!cfg || !out || (nquad > 0 && !ipqua) || (ntria > 0 && !iptri) || !bface || !esuv1 || !esuv2)
        return fail(LAG2D_EINVAL, "null argument");
    if (ntria < 0 || nquad < 0 || ntria + nquad <= 0 || npoin <= 0 || nbfac < 0) return fail(LAG2D_EINVAL, "bad sizes");
    const lag2d_config &c = *cfg;
    const bool q1 = c.ncurv == 0 && c.nfint == 1 && c.npoly == 1 && c.ndegr == 3;
    // the curved sub-cell branch (Getrhslagmc.f90:7556-7708), P2 or P1
    const bool q2 = c.ncurv == 1 && ((c.nfint == 4 && ((c.npoly == 2 && c.ndegr == 6) || (c.npoly == 1 && c.ndegr == 3))) ||
                                     (c.nfint == 5 && c.npoly == 2 && c.ndegr == 6));
    if (!(q1 || q2))
        return fail(LAG2D_EUNSUPPORTED, "supported: ncurv=0/nfint=1/P1, ncurv=1/nfint=4/P1 or P2, ncurv=1/nfint=5/P2");
    if (c.nfint == 5) {
        // the branch's face Gauss points need the 5-point Lobatto face rule
        // (ngausf = 5, implied); ncase 2, 3, 4, 13 change its boundary
        // treatment or its first cell (Getrhslagmc.f90:7717-7730,
        // Getrhslaggauss.f90:17255-17261, GetBClag.f90:556-624), and only
        // slip and fixed walls are implemented
        if (c.ncase == 2 || c.ncase == 3 || c.ncase == 4 || c.ncase == 13 || c.ncase == -13)
            return fail(LAG2D_EUNSUPPORTED, "nfint=5: ncase 2, 3, 4, 13, -13 are not implemented");
        for (int f = 0; f < nbfac; ++f)
            if (bface[5 * (size_t)f + 2] != 22 && bface[5 * (size_t)f + 2] != 25)
                return fail(LAG2D_EUNSUPPORTED, "nfint=5: slip walls (code 22) and fixed walls (25) only");
        if (rs.nrun > 1) return fail(LAG2D_EUNSUPPORTED, "nfint=5: ensembles run the nfint=4 branch");
    }
    if ((c.nrz != 0 && c.nrz != 2) || c.ndens != 1 || c.nlimi != 6 || c.nriem != 1)
        return fail(LAG2D_EUNSUPPORTED, "supported: nrz=0 or 2, ndens=1, nlimi=6, nriem=1");
    if (c.nrz == 2) {
        if (!q1) return fail(LAG2D_EUNSUPPORTED, "nrz=2: straight P1 path only (ncurv=0, nfint=1)");
        // ncase 1, 2, 12 place wall normals by node position and 3, 4 have
        // pressure boundaries (GetBClag.f90:84-165); -7 overrides the limiter
        if (c.ncase == 1 || c.ncase == 2 || c.ncase == 3 || c.ncase == 4 || c.ncase == 12 || c.ncase == -7)
            return fail(LAG2D_EUNSUPPORTED, "nrz=2: ncase 1, 2, 3, 4, 12, -7 are not implemented");
        // slip walls (22) and free boundaries (21: no force unless ncase is
        // 3 or 4, GetBClag.f90:112-127)
        for (int f = 0; f < nbfac; ++f)
            if (bface[5 * (size_t)f + 2] != 22 && bface[5 * (size_t)f + 2] != 21)
                return fail(LAG2D_EUNSUPPORTED, "nrz=2: slip walls (code 22) and free boundaries (21) only");
    }
    if (c.nstag != 2 && c.nstag != 3) return fail(LAG2D_EUNSUPPORTED, "nstag must be 2 or 3");
    // ncase 8 lowers the CFL number once the time passes 0.6
    // (Getrhslag.f90:1785); that switch is not implemented
    if (c.ncase == 8) return fail(LAG2D_EUNSUPPORTED, "ncase 8: the mid-run CFL switch (Getrhslag.f90:1785) "
                                                      "is not implemented");
    // curved six-node triangles ride on the curved P2 branch (nfint=4) with
    // the uniform initial density of the Sedov problem (ncase 7) or a
    // per-cell one (getrhoig_triacurv2, Getrhslag.f90:17622-17757)
    const bool q2h = q2 && c.nfint == 4 && c.npoly == 2 && c.ncase == 7 && rs.nrun == 1;
    if (ntria > 0 && !q1 && !q2h)
        return fail(LAG2D_EUNSUPPORTED, "triangles: straight P1 path (ncurv=0, nfint=1), or curved P2 "
                                        "(ncurv=1, nfint=4, npoly=2) with ncase 7 and one run");
    if (ntria > 0 && (c.ncase == 1 || c.ncase == 2 || c.ncase == 3 || c.ncase == -7))
        return fail(LAG2D_EUNSUPPORTED, "triangles: the boundary-node limiter overrides of ncase 1, 2, 3, -7 "
                                        "(Getrhslagmctria.f90:539-582) are not implemented");

    // the elements-around-points CSR is checked before anything else reads
    // it (the node->corner and WENO-neighbour builders index esuv1 by it)
    if (esuv2[0] != 0) return fail(LAG2D_EINVAL, "esuv2(1) must be 0 (CSR offsets)");
    for (int p = 0; p < npoin; ++p)
        if (esuv2[p + 1] < esuv2[p]) return fail(LAG2D_EINVAL, "esuv2 not monotone");

    lag2d_ctx *x = new lag2d_ctx();
    x->cfg = c;
    const int nv = c.ncurv == 0 ? 4 : 9;
    const int nc = ntria + nquad;
    x->nc = nc;
    x->ntria = ntria;
    x->np = npoin;
    x->nb = nbfac;
    x->nd = c.ndegr;
    x->nvq = nv;
    x->nmatr = c.npoly == 1 ? 4 : 16;
    int rc = 0;
    auto bad = [&](int r) {
        delete x;
        return r;
    };
    if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(LAG2D_EDEVICE, "hipStreamCreate failed"));
    const int nrun = rs.nrun;
    x->nrun = nrun;

    // ---- host-side mesh preprocessing (0-based, SoA; triangles first, their
    // unused fourth node slot repeats the third node)
    // triangle nodes per cell: 3 straight, 6 curved (corners, then the
    // mid-nodes of edges 1-2, 2-3, 3-1)
    const int nvt = c.ncurv == 1 ? 6 : 3;
    std::vector<int> ipq((size_t)nv * nc);
    for (int it = 0; it < ntria; ++it)
        for (int k = 0; k < nv; ++k) {
            int p = iptri[(size_t)(k < nvt ? k : 2) + (size_t)nvt * it] - 1;
            if (p < 0 || p >= npoin) return bad(fail(LAG2D_EINVAL, "iptri node out of range"));
            ipq[(size_t)k * nc + it] = p;
        }
    for (int iq = 0; iq < nquad; ++iq)
        for (int k = 0; k < nv; ++k) {
            int p = ipqua[(size_t)k + (size_t)nv * iq] - 1;
            if (p < 0 || p >= npoin) return bad(fail(LAG2D_EINVAL, "ipqua node out of range"));
            ipq[(size_t)k * nc + ntria + iq] = p;
        }
    const int ne = esuv2[npoin];
    if (ne != nvt * ntria + nv * nquad) return bad(fail(LAG2D_EINVAL, "esuv2(npoin+1) != nvtri*ntria + nvqua*nquad"));
    std::vector<int> e1(ne), e2(esuv2, esuv2 + npoin + 1);
    for (int k = 0; k < ne; ++k) {
        e1[k] = esuv1[k] - 1;
        if (e1[k] < 0 || e1[k] >= nc) return bad(fail(LAG2D_EINVAL, "esuv1 element out of range"));
    }
    for (int p = 0; p < npoin; ++p)
        if (e2[p + 1] < e2[p]) return bad(fail(LAG2D_EINVAL, "esuv2 not monotone"));
    // node -> corner CSR in the reference's accumulation order
    std::vector<int> n2c, n2c_off;
    if ((rc = build_node_corners(*x, ipq, e1, e2, n2c, n2c_off))) return bad(rc);
    // boundary flags and the symmetry colouring of the velocity limiter
    std::vector<uint8_t> nflag(npoin, 0), ncol(npoin, 0);
    for (int f = 0; f < nbfac; ++f) {
        const int *b = bface + 5 * (size_t)f;
        int nodes[3] = {b[0] - 1, b[1] - 1, nv == 9 ? b[4] - 1 : -1};
        for (int k = 0; k < 3; ++k)
            if (nodes[k] >= npoin) return bad(fail(LAG2D_EINVAL, "bface node out of range"));
        if (nodes[0] < 0 || nodes[1] < 0 || (nv == 9 && nodes[2] < 0))
            return bad(fail(LAG2D_EINVAL, "bface node out of range"));
        // wall projection acts on the face's two end nodes (Limiterlag.f90:608-619);
        // fixed walls exclude all face nodes from the solve (:476-481)
        // (the sub-cell solver of the curved path projects all three face
        // nodes, Getrhslaggauss.f90:4515-4545)
        const int nw = nv == 9 ? 3 : 2;
        if (b[2] == 22) {
            for (int k = 0; k < nw; ++k) nflag[nodes[k]] |= (b[3] == 221) ? 1 : (b[3] == 222 ? 2 : 0);
        } else if (b[2] == 25) {
            for (int k = 0; k < nw; ++k) nflag[nodes[k]] |= 3;
            for (int k = 0; k < (nv == 9 ? 3 : 2); ++k) nflag[nodes[k]] |= 4;
        }
    }
    if (c.nfint == 5)  // cell-centre nodes: u* = 0 after the solve (Getrhslaggauss.f90:17475)
        for (int ie = 0; ie < nc; ++ie) nflag[ipq[(size_t)8 * nc + ie]] |= 16;
    if (c.ncase != 1) {  // Limiterlag.f90:8212-8238
        const int nf = nv == 9 ? 3 : 2;
        for (int f = 0; f < nbfac; ++f) {
            const int *b = bface + 5 * (size_t)f;
            int nodes[3] = {b[0] - 1, b[1] - 1, nv == 9 ? b[4] - 1 : 0};
            if (b[2] == 22 && b[3] == 221)
                for (int k = 0; k < nf; ++k) ncol[nodes[k]] = 3;
        }
        for (int f = 0; f < nbfac; ++f) {
            const int *b = bface + 5 * (size_t)f;
            int nodes[3] = {b[0] - 1, b[1] - 1, nv == 9 ? b[4] - 1 : 0};
            if (b[2] == 22 && b[3] == 222)
                for (int k = 0; k < nf; ++k) ncol[nodes[k]] = (ncol[nodes[k]] == 3) ? 10 : 2;
        }
    }

    // RZ slip-wall forces: the (at most two) wall faces of each node, in
    // face order, as other node*4 + (1: subcode 221, 2: 222, 0: other)
    std::vector<int> wall;
    if (c.nrz == 2) {
        wall.assign((size_t)2 * npoin, -1);
        for (int f = 0; f < nbfac; ++f) {
            const int *b = bface + 5 * (size_t)f;
            if (b[2] != 22) continue;
            const int typ = b[3] == 221 ? 1 : (b[3] == 222 ? 2 : 0);
            for (int k = 0; k < 2; ++k) {
                const int p = b[k] - 1, o = b[1 - k] - 1;
                nflag[p] |= 8;
                const int j = wall[p] < 0 ? 0 : 1;
                if (j == 1 && wall[(size_t)npoin + p] >= 0)
                    return bad(fail(LAG2D_EINVAL, "nrz=2: a node on more than two wall faces"));
                wall[(size_t)j * npoin + p] = o * 4 + typ;
            }
        }
    }

    Dev &d = x->d;
    memset(&d, 0, sizeof(d));
    d.nrz = c.nrz;
    d.alfrz = (c.nrz == 1 || c.nrz == 2) ? 1.0 : 0.0;
    d.nc = nc; d.ntria = ntria; d.np = npoin; d.nb = nbfac; d.nd = c.ndegr; d.nvq = nv; d.nmatr = x->nmatr;
    d.nstag = c.nstag; d.ncase = c.ncase; d.ncurv = c.ncurv; d.npoly = c.npoly; d.ndt = c.ndt;
    d.gamlg = c.gamlg; d.cfl = c.cfl; d.cimpd = c.cimpd; d.cdrho = c.cdrho; d.tend = c.tend;
    d.slpdu = (c.gamlg + 1.0) / 2.0;
    int *t_ipq, *t_e1, *t_e2, *t_n2c, *t_n2co;
    uint8_t *t_fl, *t_col;
    const size_t nU = (size_t)4 * c.ndegr * nc;
    x->nblk_dt = nrun == 1 ? std::min(nblocks(nc), 1024) : 4;
    if (q2) x->nrec = Q2_REC;
    if ((rc = x->alloc(&t_ipq, ipq.size())) || (rc = x->alloc(&t_e1, e1.size())) ||
        (rc = x->alloc(&t_e2, e2.size())) || (rc = x->alloc(&t_n2c, n2c.size())) ||
        (rc = x->alloc(&t_n2co, n2c_off.size())) || (rc = x->alloc(&t_fl, (size_t)npoin)) ||
        (rc = x->alloc(&t_col, (size_t)npoin)) || (rc = x->alloc(&d.xn, (size_t)npoin)) ||
        (rc = x->alloc(&d.yn, (size_t)npoin)) || (rc = x->alloc(&d.xs, (size_t)npoin)) ||
        (rc = x->alloc(&d.ys, (size_t)npoin)) || (rc = x->alloc(&x->x0, (size_t)2 * npoin)) ||
        (rc = x->alloc(&d.Un, nU)) || (rc = x->alloc(&d.Us, nU)) || (rc = x->alloc(&d.R, nU)) ||
        (rc = x->alloc(&x->A, (size_t)x->nmatr * nc)) || (rc = x->alloc(&d.G, (size_t)24 * nc)) ||
        (rc = x->alloc(&d.usx, (size_t)npoin)) || (rc = x->alloc(&d.usy, (size_t)npoin)) ||
        (rc = x->alloc(&d.cavg, (size_t)5 * nc)) || (rc = x->alloc(&d.frame, (size_t)4 * nc)) ||
        (rc = x->alloc(&d.nmm, (size_t)10 * npoin)) || (rc = x->alloc(&d.lim, (size_t)9 * nc)) ||
        (rc = x->alloc(&d.rec, (size_t)x->nrec * (q2 ? (nc + 63) / 64 * 64 : nc))) ||
        (rc = x->alloc(&d.red, (size_t)2 * nrun * x->nblk_dt)) || (rc = x->alloc(&d.scal, (size_t)2 * nrun + 2)))
        return bad(rc);
    d.nrun = nrun;
    if (nrun > 1) {
        // run index of every cell and node
        std::vector<int> runc(nc), runn(npoin);
        for (int r = 0; r < nrun; ++r) {
            for (int ie = rs.coff[r]; ie < rs.coff[r + 1]; ++ie) runc[ie] = r;
            for (int p = rs.noff[r]; p < rs.noff[r + 1]; ++p) runn[p] = r;
        }
        for (int ie = 0; ie < nc; ++ie)
            for (int k = 0; k < nv; ++k) {
                const int p = ipq[(size_t)k * nc + ie];
                if (runn[p] != runc[ie]) return bad(fail(LAG2D_EINVAL, "a cell uses a node of another run"));
            }
        int *t_runc, *t_runn, *t_coff, *t_act;
        double *t_tend;
        if ((rc = x->alloc(&t_runc, (size_t)nc)) || (rc = x->alloc(&t_runn, (size_t)npoin)) ||
            (rc = x->alloc(&t_coff, (size_t)nrun + 1)) || (rc = x->alloc(&t_act, (size_t)nrun)) ||
            (rc = x->alloc(&t_tend, (size_t)nrun)))
            return bad(rc);
        std::vector<int> ones(nrun, 1);
        if (hipMemcpy(t_runc, runc.data(), runc.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t_runn, runn.data(), runn.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t_coff, rs.coff.data(), rs.coff.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t_act, ones.data(), ones.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t_tend, rs.tend.data(), rs.tend.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return bad(fail(LAG2D_EDEVICE, "ensemble table upload"));
        d.runc = t_runc; d.runn = t_runn; d.coff = t_coff; d.act = t_act; d.tendv = t_tend;
        x->tend_host = rs.tend;
    }
    if (q2) {
        // face neighbours for the WENO stencils: neighbour id, or nc + f on
        // a boundary face (the reference keeps a ghost-cell id there)
        std::vector<int> esq((size_t)4 * nc);
        for (int ie = 0; ie < nc; ++ie)
            for (int f = 0; f < 4; ++f) {
                int a = ipq[(size_t)f * nc + ie], b = ipq[(size_t)((f + 1) & 3) * nc + ie];
                int nb = nc + f;
                for (int k = e2[a]; k < e2[a + 1] && nb >= nc; ++k) {
                    int je = e1[k];
                    if (je == ie) continue;
                    for (int l = e2[b]; l < e2[b + 1]; ++l)
                        if (e1[l] == je) {
                            nb = je;
                            break;
                        }
                }
                esq[(size_t)f * nc + ie] = nb;
            }
        int *t_esq;
        if ((rc = x->alloc(&d.msub, (size_t)4 * nc)) || (rc = x->alloc(&d.Ul, nU)) ||
            (rc = x->alloc(&d.gq, (size_t)48 * nc)) || (rc = x->alloc(&d.flag, (size_t)nc)) ||
            (rc = x->alloc(&d.tlist, (size_t)nc)) || (rc = x->alloc(&d.tcount, 2)) ||
            (rc = x->alloc(&d.tsum, 1)) ||
            (rc = x->alloc(&t_esq, esq.size())))
            return bad(rc);
        d.esqua = t_esq;
        d.wenobad = d.tcount + 1;
        d.weno_strict = c.strict;
        if (hipMemcpy(t_esq, esq.data(), esq.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(d.tcount, 0, 2 * sizeof(int)) != hipSuccess ||
            hipMemset(d.tsum, 0, sizeof(unsigned long long)) != hipSuccess ||
            hipMemset(d.flag, 0, (size_t)nc) != hipSuccess)
            return bad(fail(LAG2D_EDEVICE, "curved-path table upload"));
        d.nfint = c.nfint;
        if (c.nfint == 5) {
            // the other side of every cell face: the neighbour sharing the
            // face's mid-node, or the boundary face holding it
            std::vector<int> fnb((size_t)4 * nc, -1), mid(npoin, -1);
            for (int ie = 0; ie < nc; ++ie)
                for (int lf = 0; lf < 4; ++lf) {
                    const int m = ipq[(size_t)(4 + lf) * nc + ie], me = 4 * ie + lf;
                    if (mid[m] < 0) {
                        mid[m] = me;
                        continue;
                    }
                    const int o = mid[m];
                    if (fnb[(size_t)(o & 3) * nc + (o >> 2)] != -1)
                        return bad(fail(LAG2D_EINVAL, "nfint=5: a mid-node shared by more than two cells"));
                    fnb[(size_t)lf * nc + ie] = o;
                    fnb[(size_t)(o & 3) * nc + (o >> 2)] = me;
                }
            for (int f = 0; f < nbfac; ++f) {
                const int *b = bface + 5 * (size_t)f;
                const int o = mid[b[4] - 1];
                if (o < 0 || fnb[(size_t)(o & 3) * nc + (o >> 2)] >= 0)
                    return bad(fail(LAG2D_EINVAL, "nfint=5: a boundary face's mid-node is not on a single cell face"));
                const int bc = b[2] == 25 ? 3 : (b[3] == 221 ? 1 : (b[3] == 222 ? 2 : 0));
                const int rev = (b[0] - 1) != ipq[(size_t)(o & 3) * nc + (o >> 2)];
                fnb[(size_t)(o & 3) * nc + (o >> 2)] = -(1 + bc + 4 * rev);
            }
            int *t_fnb;
            if ((rc = x->alloc(&d.drho, (size_t)4 * nc)) || (rc = x->alloc(&d.fg, (size_t)48 * nc)) ||
                (rc = x->alloc(&t_fnb, fnb.size())))
                return bad(rc);
            d.fnb = t_fnb;
            if (hipMemcpy(t_fnb, fnb.data(), fnb.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
                return bad(fail(LAG2D_EDEVICE, "face table upload"));
        }
    }
    if (c.nrz == 2) {
        int *t_wall;
        if ((rc = x->alloc(&t_wall, wall.size()))) return bad(rc);
        if (hipMemcpy(t_wall, wall.data(), wall.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return bad(fail(LAG2D_EDEVICE, "wall table upload"));
        d.wall = t_wall;
    }
    d.ipq = t_ipq; d.esuv1 = t_e1; d.esuv2 = t_e2; d.n2c = t_n2c; d.n2c_off = t_n2co;
    d.nflag = t_fl; d.ncol = t_col; d.A = x->A;
    d.x0 = x->x0; d.y0 = x->x0 + npoin;
    hipError_t e = hipSuccess;
    e = hipMemcpy(t_ipq, ipq.data(), ipq.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_e1, e1.data(), e1.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_e2, e2.data(), e2.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_n2c, n2c.data(), n2c.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_n2co, n2c_off.data(), n2c_off.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_fl, nflag.data(), npoin, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t_col, ncol.data(), npoin, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bad(fail(LAG2D_EDEVICE, std::string("mesh upload: ") + hipGetErrorString(e)));
    std::vector<double> sc((size_t)2 * nrun + 2, 0.0);
    for (int r = 0; r < nrun; ++r) {
        sc[2 * r] = nrun == 1 ? c.dtfix : rs.dtfix[r];
        sc[2 * r + 1] = nrun == 1 ? c.tacru : rs.tacru[r];
    }
    if (hipMemcpy(d.scal, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bad(fail(LAG2D_EDEVICE, "scalar upload"));
    x->ipq_host = std::move(ipq);
    if (const char *v = std::getenv("LAG2D_CORNER_VARIANT")) {
        const int k = std::atoi(v);
        if (k >= 1 && k <= LAG2D_CORNER_VARIANTS) x->corner_variant = k;
    }
    // the update reads each cell's M^-1 and residual once where the state
    // outgrows the caches (LAG2D_UPDATE_VARIANT = 1 per DOF, 2 per cell)
    x->update_cells = nc >= (1 << 18);
    if (const char *v = std::getenv("LAG2D_UPDATE_VARIANT")) {
        const int k = std::atoi(v);
        if (k == 1 || k == 2) x->update_cells = k == 2;
    }
    // cell volume and Persson indicator: a thread per cell where the mesh
    // fills the chip with it (LAG2D_GEOM_VARIANT = 1 8 lanes per cell, 2 a
    // thread per cell)
    x->geom_cells = nc >= GEOM_CELLS_MIN;
    if (const char *v = std::getenv("LAG2D_GEOM_VARIANT")) {
        const int k = std::atoi(v);
        if (k == 1 || k == 2) x->geom_cells = k == 2;
    }
    if (const char *v = std::getenv("LAG2D_RHS_VARIANT")) {
        const int k = std::atoi(v);
        if (k >= 1 && k <= LAG2D_RHS_VARIANTS) x->rhs_variant = k;
    }
    if (q2 && c.npoly == 1) {  // curved P1 has one corner and one cell-RHS form
        x->corner_variant = 8;
        x->rhs_variant = 2;
    }
    if (q2 && (c.nfint == 5 || ntria > 0)) {
        // the nfint=5 branch and hybrid meshes: a lane per sub-cell corner,
        // a thread per cell RHS (the forms that take their options); on the
        // nfint=5 branch two waves per sub-cell corner and 8 cells per
        // workgroup in the cell RHS where the mesh does not fill the chip
        // with a thread per cell
        const bool small5 = c.nfint == 5 && ntria == 0 && !x->geom_cells;
        x->corner_variant = small5 ? 9 : 2;
        x->rhs_variant = small5 ? 1 : 2;
    }
    *out = x;
    return 0;
}

extern "C" int lag2d_create(const lag2d_config *cfg, int nquad, int npoin, int nbfac, const int *ipqua,
                            const int *bface, const int *esuv1, const int *esuv2, lag2d_ctx **out)
{
    return create_impl(cfg, 0, nquad, npoin, nbfac, nullptr, ipqua, bface, esuv1, esuv2, RunSpec(), out);
}

extern "C" int lag2d_create_hybrid(const lag2d_config *cfg, int ntria, int nquad, int npoin, int nbfac,
                                   const int *iptri, const int *ipqua, const int *bface, const int *esuv1,
                                   const int *esuv2, lag2d_ctx **out)
{
    return create_impl(cfg, ntria, nquad, npoin, nbfac, iptri, ipqua, bface, esuv1, esuv2, RunSpec(), out);
}

extern "C" int lag2d_create_ensemble(const lag2d_config *cfg, int nrun, const int *cell_off, const int *node_off,
                                     const double *tend, const double *dtfix, const double *tacru, int nquad,
                                     int npoin, int nbfac, const int *ipqua, const int *bface, const int *esuv1,
                                     const int *esuv2, lag2d_ctx **out)
{
    if (!cfg || !cell_off || !node_off || !tend || !dtfix || !tacru) return fail(LAG2D_EINVAL, "null argument");
    if (nrun < 1) return fail(LAG2D_EINVAL, "nrun < 1");
    if (cfg->ncurv != 1 || cfg->npoly != 2)
        return fail(LAG2D_EUNSUPPORTED, "ensembles run the curved P2 path (ncurv=1, npoly=2)");
    if (cell_off[0] != 0 || cell_off[nrun] != nquad || node_off[0] != 0 || node_off[nrun] != npoin)
        return fail(LAG2D_EINVAL, "run offsets must start at 0 and end at nquad / npoin");
    RunSpec rs;
    rs.nrun = nrun;
    for (int r = 0; r < nrun; ++r) {
        if (cell_off[r + 1] <= cell_off[r] || node_off[r + 1] <= node_off[r])
            return fail(LAG2D_EINVAL, "empty run or decreasing offsets");
        if (!(tend[r] > 0.0)) return fail(LAG2D_EINVAL, "tend must be positive");
    }
    rs.coff.assign(cell_off, cell_off + nrun + 1);
    rs.noff.assign(node_off, node_off + nrun + 1);
    rs.tend.assign(tend, tend + nrun);
    rs.dtfix.assign(dtfix, dtfix + nrun);
    rs.tacru.assign(tacru, tacru + nrun);
    return create_impl(cfg, 0, nquad, npoin, nbfac, nullptr, ipqua, bface, esuv1, esuv2, rs, out);
}

extern "C" int lag2d_upload(lag2d_ctx *x, const double *coord, const double *coold, const double *geoel,
                            const double *amatr, const double *unkno, const double *ustar)
{
    if (!x || !coord || !coold || !geoel || !amatr || !unkno || !ustar) return fail(LAG2D_EINVAL, "null argument");
    const int nc = x->nc, np = x->np, nd = x->nd, nm = x->nmatr;
    Dev &d = x->d;
    std::vector<double> h((size_t)std::max({2 * np, 4 * nd * nc, 24 * nc, nm * nc}));
    auto up = [&](double *dst, size_t n) { return hipMemcpy(dst, h.data(), n * sizeof(double), hipMemcpyHostToDevice); };
    for (int p = 0; p < np; ++p) h[p] = coord[2 * (size_t)p];
    HIPCHK(up(d.xn, np));
    for (int p = 0; p < np; ++p) h[p] = coord[2 * (size_t)p + 1];
    HIPCHK(up(d.yn, np));
    for (int p = 0; p < np; ++p) {
        h[p] = coold[2 * (size_t)p];
        h[np + p] = coold[2 * (size_t)p + 1];
    }
    HIPCHK(up(x->x0, 2 * (size_t)np));
    for (int p = 0; p < np; ++p) h[p] = ustar[2 * (size_t)p];
    HIPCHK(up(d.usx, np));
    for (int p = 0; p < np; ++p) h[p] = ustar[2 * (size_t)p + 1];
    HIPCHK(up(d.usy, np));
    for (int ie = 0; ie < nc; ++ie)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < nd; ++id)
                h[((size_t)q * nd + id) * nc + ie] = unkno[id + (size_t)nd * (q + 4 * (size_t)ie)];
    HIPCHK(up(d.Un, (size_t)4 * nd * nc));
    for (int ie = 0; ie < nc; ++ie)
        for (int k = 0; k < 24; ++k) h[(size_t)k * nc + ie] = geoel[k + 24 * (size_t)ie];
    HIPCHK(up(d.G, (size_t)24 * nc));
    x->geo_host.assign(geoel, geoel + (size_t)24 * nc);
    for (int ie = 0; ie < nc; ++ie)
        for (int k = 0; k < nm; ++k) h[(size_t)k * nc + ie] = amatr[k + (size_t)nm * ie];
    HIPCHK(up(x->A, (size_t)nm * nc));
    launch_cellvol(x, d.xn, d.yn, 0, false);
    HIPCHK(hipGetLastError());
    int rc = upload_extra(*x);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

// Driver-side initialisation on the device (curved P2 path on quads,
// straight P1 path planar or RZ): the arrays the reference's start-up
// computes before the time loop -- geometry, inverse mass matrices, DOFs --
// from the coordinates and a piecewise constant state per cell, in place of
// an upload of host-computed arrays.
extern "C" int lag2d_init_state(lag2d_ctx *x, const double *coord, const double *rho, const double *u,
                                const double *v, const double *p, int nblast, const int *blast,
                                const double *energy)
{
    if (!x || !coord || !rho || !u || !v || !p || (nblast > 0 && (!blast || !energy)) || nblast < 0)
        return fail(LAG2D_EINVAL, "null or negative argument");
    const lag2d_config &c = x->cfg;
    const bool p1 = c.ncurv == 0 && c.npoly == 1 && x->nd == 3;
    if (!p1 && (c.ncurv != 1 || c.npoly != 2 || c.nrz != 0 || x->ntria > 0))
        return fail(LAG2D_EUNSUPPORTED, "device initialisation covers the planar curved P2 path (ncurv=1, npoly=2) "
                                        "on quads and the straight P1 path (ncurv=0, npoly=1)");
    const int nc = x->nc, np = x->np;
    for (int k = 0; k < nblast; ++k)
        if (blast[k] < 1 || blast[k] > nc) return fail(LAG2D_EINVAL, "blast cell out of range (1-based)");
    for (int ie = 0; ie < nc; ++ie)
        if (!(rho[ie] > 0.0)) return fail(LAG2D_EINVAL, "initial density must be positive");
    Dev &d = x->d;
    std::vector<double> h((size_t)std::max(2 * np, 6 * nc), 0.0);
    auto up = [&](double *dst, size_t n) { return hipMemcpy(dst, h.data(), n * sizeof(double), hipMemcpyHostToDevice); };
    for (int q = 0; q < 2; ++q) {
        for (int k = 0; k < np; ++k) h[k] = coord[2 * (size_t)k + q];
        HIPCHK(up(q == 0 ? d.xn : d.yn, np));
    }
    for (int k = 0; k < np; ++k) {
        h[k] = coord[2 * (size_t)k];
        h[np + k] = coord[2 * (size_t)k + 1];
    }
    HIPCHK(up(x->x0, 2 * (size_t)np));  // coold = coord
    HIPCHK(hipMemset(d.usx, 0, np * sizeof(double)));
    HIPCHK(hipMemset(d.usy, 0, np * sizeof(double)));
    // the per-cell state, staged in the residual buffer (rewritten by the first step)
    std::fill(h.begin(), h.end(), 0.0);
    for (int ie = 0; ie < nc; ++ie) {
        h[ie] = rho[ie];
        h[nc + ie] = u[ie];
        h[2 * (size_t)nc + ie] = v[ie];
        h[3 * (size_t)nc + ie] = p[ie];
    }
    for (int k = 0; k < nblast; ++k) {
        h[4 * (size_t)nc + blast[k] - 1] = energy[k];
        h[5 * (size_t)nc + blast[k] - 1] = 1.0;
    }
    HIPCHK(up(d.R, 6 * (size_t)nc));
    if (p1) {
        if (c.nrz == 2)
            hipLaunchKernelGGL(k_init_p1<1>, dim3(nblocks(nc)), dim3(LAG2D_BLOCK), 0, x->stream, d,
                               (const double *)d.R, x->A);
        else
            hipLaunchKernelGGL(k_init_p1<0>, dim3(nblocks(nc)), dim3(LAG2D_BLOCK), 0, x->stream, d,
                               (const double *)d.R, x->A);
    } else {
        hipLaunchKernelGGL(k_init_q2, dim3(nblocks(nc)), dim3(LAG2D_BLOCK), 0, x->stream, d, (const double *)d.R,
                           x->A);
    }
    HIPCHK(hipGetLastError());
    // the initial density of the curved path's sub-cell masses
    if (!p1 && !d.rhoi) {
        double *t;
        int rc = x->alloc(&t, (size_t)nc);
        if (rc) return rc;
        d.rhoi = t;
    }
    if (!p1)
        HIPCHK(hipMemcpyAsync((void *)d.rhoi, rho, (size_t)nc * sizeof(double), hipMemcpyHostToDevice, x->stream));
    x->geo_host.resize((size_t)24 * nc);
    std::vector<double> g((size_t)24 * nc);
    HIPCHK(hipMemcpyAsync(g.data(), d.G, g.size() * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int ie = 0; ie < nc; ++ie)
        for (int k = 0; k < 24; ++k) x->geo_host[k + 24 * (size_t)ie] = g[(size_t)k * nc + ie];
    launch_cellvol(x, d.xn, d.yn, 0, false);
    HIPCHK(hipGetLastError());
    int rc = upload_extra(*x);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

// ---- launch helpers with optional per-kernel event timing
void lag2d_ctx::tic(int k)
{
    if (!timing) return;
    auto &t = timers[k];
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, stream);
    t.ev.push_back({a, b});
}

void lag2d_ctx::toc(int k)
{
    if (!timing) return;
    hipEventRecord(timers[k].ev.back().second, stream);
}

#define LAUNCH(kid, kern, grid, ...)                                                   \
    do {                                                                               \
        x->tic(kid);                                                                   \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(LAG2D_BLOCK), 0, x->stream, __VA_ARGS__); \
        x->toc(kid);                                                                   \
    } while (0)

const char *const lag2d_kernel_names[KID_COUNT] = {
    "k_cellavg_q1", "k_nodeminmax", "k_limit_corner_q1", "k_node_q1", "k_cellrhs_q1", "k_dt_partial",
    "k_dt_final",   "k_update",     "k_nodemove",        "k_cellvol",
    "k_q2_prep",    "k_q2_weno",    "k_q2_corner",       "k_node_q2", "k_cellrhs_q2", "k_run_flags",
    "k_limit_corner_rz", "k_node_rz", "k_cellrhs_rz",    "k_schedule",
    "k_adj_cellrhs", "k_adj_corner", "k_adj_node+gather+nodemove", "k_limit_corner_q1<frozen>", "k_q2g_face"};

static const double RKLAG[2][3] = {{0.0, 0.5, 0.0}, {0.0, 0.75, 0.3333333333333333}};

// Name of kernel id k as timed on this context: the curved path's corner and
// cell-RHS slots name the variant the autotuner kept.
static const char *kernel_name(const lag2d_ctx *x, int k)
{
    if (k == KID_Q2_CORNER && x->corner_variant == 9) return "k_q2_corner_wu2";
    if (k == KID_Q2_CORNER && x->corner_variant >= 7) return "k_q2_corner_wu";
    if (k == KID_Q2_CORNER && x->corner_variant >= 2) return "k_q2_corner_sc";
    if (k == KID_CELLRHS_Q2 && x->rhs_variant >= 2) return "k_cellrhs_q2_tpc";
    if (k == KID_Q2_WENO && x->nd == 3) return "k_c1_weno";
    if (x->adj.curved && k == KID_ADJ_CELLRHS) return "k_adjc_cellrhs";
    if (x->adj.curved && k == KID_ADJ_CORNER) return "k_adjc_corner";
    if (x->adj.curved && k == KID_ADJ_NODES) return "k_adjc_move+node+gather";
    return lag2d_kernel_names[k];
}

// The dtfix and epsaw that getsolvdg_lag sets at the start of step itime
// (Getrhslag.f90:1663-1737) for the problems that fix them; returns whether
// dt is fixed.  epsaw (the RZ basis-motion lead) is 0.1 otherwise.
static bool schedule(const lag2d_config &c, int itime, double *dt, double *epsaw)
{
    *epsaw = 1.e-1;
    if (c.nrz >= 1 && c.ncase == 7) {
        if (itime <= 2000) { *dt = 1e-6; *epsaw = 1.0; }
        else if (itime <= 6000) { *dt = 2.5e-6; *epsaw = 4.e-1; }
        else if (itime <= 10000) { *dt = 5.e-6; *epsaw = 2.e-1; }
        else { *dt = 5.e-5; *epsaw = 2.e-2; }
        return true;
    }
    if (c.nrz >= 1 && c.ncase == 11) {
        if (itime <= 7500) { *dt = 1.e-4; *epsaw = 1.e-3; }
        else { *dt = 2.5e-6; *epsaw = 2.e-2; }
        return true;
    }
    if (c.ncase == -4) { *dt = 2.e-6; return true; }
    if (c.ncase == 13) {
        *dt = itime <= 7500 ? 1.e-4 : (itime <= 17500 ? 2.e-5 : 1e-5);
        return true;
    }
    if (c.ncase == -14) {
        if (itime <= 1000) { *dt = 5.e-4; *epsaw = 1.e-2; }
        else *dt = 1.e-4;
        return true;
    }
    return false;
}

static void launch_cellvol(lag2d_ctx *x, double *xs, double *ys, int set_centre, bool timed)
{
    const int grid = (x->nc + CV_CELLS - 1) / CV_CELLS;
    if (timed) x->tic(KID_CELLVOL);
    const int rz = x->cfg.nrz == 2;
    if (x->geom_cells) {
        auto k = rz ? k_cellvol_tpc<1> : k_cellvol_tpc<0>;
        hipLaunchKernelGGL(k, dim3(nblocks(x->nc)), dim3(LAG2D_BLOCK), 0, x->stream, x->d, xs, ys, set_centre);
    } else {
        auto k = rz ? k_cellvol<1> : k_cellvol<0>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(LAG2D_BLOCK), 0, x->stream, x->d, xs, ys, set_centre);
    }
    if (timed) x->toc(KID_CELLVOL);
}

// RHS of one stage (0-based) of step itime on state Uin and stage
// coordinates (ds.xs/ds.ys); leaves the limited DOFs in Us.  apply_update
// fuses the M^-1 solve and RK blend.
static int launch_rhs(lag2d_ctx *x, Dev &ds, const double *Uin, double alfa, int apply_update, int itime,
                      int stage)
{
    const int nc = x->nc, np = x->np;
    if (x->cfg.nrz == 2) {
        double dt, epsaw;
        schedule(x->cfg, itime, &dt, &epsaw);
        // the stage's mass centre and mass matrix are recomputed from the
        // second stage of the first step on (Getrhslagmc.f90:6882-6885)
        const int rev = (itime > 1 || stage > 0) ? 1 : 0;
        LAUNCH(KID_CELLAVG_Q1, k_cellavg_q1, nblocks(nc), ds, Uin);
        LAUNCH(KID_NODEMINMAX, k_nodeminmax, nblocks(np), ds);
        LAUNCH(KID_LIMIT_RZ, k_limit_corner_rz, nblocks(nc), ds, Uin, rev);
        LAUNCH(KID_NODE_RZ, k_node_rz, nblocks(np), ds);
        LAUNCH(KID_CELLRHS_RZ, k_cellrhs_rz, nblocks(nc), ds, Uin, alfa, apply_update, epsaw);
    } else if (x->cfg.ncurv == 0) {
        LAUNCH(KID_CELLAVG_Q1, k_cellavg_q1, nblocks(nc), ds, Uin);
        LAUNCH(KID_NODEMINMAX, k_nodeminmax, nblocks(np), ds);
        LAUNCH(KID_LIMIT_Q1, k_limit_corner_q1<0>, nblocks(nc), ds, Uin);
        LAUNCH(KID_NODE_Q1, k_node_q1, nblocks(np), ds);
        LAUNCH(KID_CELLRHS_Q1, k_cellrhs_q1, nblocks(nc), ds, Uin, alfa, apply_update);
    } else {
        int rc = launch_rhs_q2(x, ds, Uin, alfa, apply_update);
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static int step_impl(lag2d_ctx *x, int itime)
{
    Dev &d = x->d;
    const int nc = x->nc, np = x->np;
    const int ns = x->cfg.nstag;
    if (x->nrun > 1) LAUNCH(KID_RUN_FLAGS, k_run_flags, nblocks(x->nrun), d);
    double dtfix, epsaw;
    if (schedule(x->cfg, itime, &dtfix, &epsaw)) LAUNCH(KID_SCHEDULE, k_schedule, nblocks(x->nrun), d, dtfix);
    for (int s = 0; s < ns; ++s) {
        const double alfa = RKLAG[ns - 2][s];
        Dev ds = d;
        if (s == 0) {
            ds.xs = d.xn;
            ds.ys = d.yn;
        }
        const double *Uin = (s == 0) ? d.Un : d.Us;
        int rc = launch_rhs(x, ds, Uin, alfa, s > 0, itime, s);
        if (rc) return rc;
        if (s == 0) {
            LAUNCH(KID_DT_PARTIAL, k_dt_partial, x->nrun * x->nblk_dt, ds, Uin, x->nblk_dt);
            LAUNCH(KID_DT_FINAL, k_dt_final, x->nrun, d, itime, x->nblk_dt);
            if (x->update_cells)
                LAUNCH(KID_UPDATE, k_update_cells, nblocks(4 * nc), d, alfa);
            else
                LAUNCH(KID_UPDATE, k_update, nblocks(4 * x->nd * nc), d, alfa);
        }
        LAUNCH(KID_NODEMOVE, k_nodemove, nblocks(np), d, alfa, s == 0 ? d.xn : d.xs, s == 0 ? d.yn : d.ys);
        // the volumes of the stage coordinates: the curved path's centre
        // nodes, and the RZ path's mean radii, are read by the next stage
        if ((x->cfg.ncurv == 1 || x->cfg.nrz == 2) && s + 1 < ns) launch_cellvol(x, d.xs, d.ys, x->cfg.ncurv == 1, true);
    }
    // final volumes; curved cells reset their centre node here for the last stage
    launch_cellvol(x, d.xs, d.ys, x->cfg.ncurv == 1, true);
    HIPCHK(hipGetLastError());
    std::swap(d.Un, d.Us);
    std::swap(d.xn, d.xs);
    std::swap(d.yn, d.ys);
    x->parity ^= 1;
    return 0;
}

// ---- discrete adjoint (straight P1 planar quads, SSP-RK2)

enum { ADJ_RECORD = 0, ADJ_REPLAY = 1, ADJ_SEGMENT = 2 };

// Segment slot layout: U^n, U^(1) [12 nc each]; x^n, x^(1), u*_prev, u*_0,
// u*_1 [2 np each]; corner records of both stages [40 nc each].
static size_t seg_stride(const lag2d_ctx *x) { return (size_t)104 * x->nc + (size_t)10 * x->np; }

// One step of the recorded forward (RECORD: the plain step, whose limiter
// coefficients and dt are written to the tape), of the frozen replay
// (REPLAY: limiter and dt read from the tape), or of a recomputed segment
// (SEGMENT: as REPLAY, keeping the stage data of slot k for the reverse).
static int adj_fwd_step(lag2d_ctx *x, int itime, int n, int mode, int k)
{
    Dev &d = x->d;
    lag2d_adj &A = x->adj;
    const int nc = x->nc, np = x->np;
    const size_t nU = (size_t)12 * nc, lim = (size_t)17 * nc;
    double *slot = A.seg + (size_t)k * A.seg_stride;
    auto cp = [&](double *dst, const double *src, size_t n) {
        return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, x->stream);
    };
    for (int s = 0; s < 2; ++s) {
        const double alfa = RKLAG[0][s];
        Dev ds = d;
        if (s == 0) {
            ds.xs = d.xn;
            ds.ys = d.yn;
        }
        const double *Uin = (s == 0) ? d.Un : d.Us;
        if (mode == ADJ_SEGMENT) {
            if (s == 0) {
                HIPCHK(cp(slot, d.Un, nU));
                HIPCHK(cp(slot + 2 * nU, d.xn, np));
                HIPCHK(cp(slot + 2 * nU + np, d.yn, np));
                HIPCHK(cp(slot + 2 * nU + 4 * (size_t)np, d.usx, np));
                HIPCHK(cp(slot + 2 * nU + 5 * (size_t)np, d.usy, np));
            } else {
                HIPCHK(cp(slot + nU, d.Us, nU));
                HIPCHK(cp(slot + 2 * nU + 2 * (size_t)np, d.xs, np));
                HIPCHK(cp(slot + 2 * nU + 3 * (size_t)np, d.ys, np));
            }
        }
        // the stage's limiter coefficients and impedance |du.n| live in the
        // tape: written by the recorded pass, read by the frozen ones
        ds.lim = A.lim_tape + ((size_t)n * 2 + s) * lim;
        ds.dlt = ds.lim + 9 * (size_t)nc;
        if (mode == ADJ_RECORD) {
            LAUNCH(KID_CELLAVG_Q1, k_cellavg_q1, nblocks(nc), ds, Uin);
            LAUNCH(KID_NODEMINMAX, k_nodeminmax, nblocks(np), ds);
            LAUNCH(KID_LIMIT_Q1, k_limit_corner_q1<0>, nblocks(nc), ds, Uin);
        } else {
            LAUNCH(KID_LIMIT_FROZEN, k_limit_corner_q1<1>, nblocks(nc), ds, Uin);
        }
        if (mode == ADJ_SEGMENT) HIPCHK(cp(slot + 2 * nU + 10 * (size_t)np + (size_t)s * 40 * nc, d.rec, (size_t)40 * nc));
        LAUNCH(KID_NODE_Q1, k_node_q1, nblocks(np), ds);
        if (mode == ADJ_SEGMENT) {
            HIPCHK(cp(slot + 2 * nU + (6 + 2 * s) * (size_t)np, d.usx, np));
            HIPCHK(cp(slot + 2 * nU + (7 + 2 * s) * (size_t)np, d.usy, np));
        }
        LAUNCH(KID_CELLRHS_Q1, k_cellrhs_q1, nblocks(nc), ds, Uin, alfa, s > 0);
        if (s == 0) {
            if (mode == ADJ_RECORD) {
                LAUNCH(KID_DT_PARTIAL, k_dt_partial, x->nblk_dt, ds, Uin, x->nblk_dt);
                LAUNCH(KID_DT_FINAL, k_dt_final, 1, d, itime, x->nblk_dt);
                HIPCHK(cp(A.dt_tape + n, d.scal, 1));
            } else {
                LAUNCH(KID_DT_FINAL, k_replay_dt, 1, d, A.dt_tape + n);
            }
            if (x->update_cells)
                LAUNCH(KID_UPDATE, k_update_cells, nblocks(4 * nc), d, alfa);
            else
                LAUNCH(KID_UPDATE, k_update, nblocks(4 * x->nd * nc), d, alfa);
        }
        LAUNCH(KID_NODEMOVE, k_nodemove, nblocks(np), d, alfa, s == 0 ? d.xn : d.xs, s == 0 ? d.yn : d.ys);
    }
    if (mode == ADJ_RECORD) launch_cellvol(x, d.xs, d.ys, 0, true);
    HIPCHK(hipGetLastError());
    std::swap(d.Un, d.Us);
    std::swap(d.xn, d.xs);
    std::swap(d.yn, d.ys);
    x->parity ^= 1;
    return 0;
}

static bool adj_supported(const lag2d_ctx *x)
{
    const lag2d_config &c = x->cfg;
    return c.ncurv == 0 && c.npoly == 1 && c.nrz == 0 && c.nstag == 2 && x->ntria == 0 && x->nrun == 1;
}

// ---- discrete adjoint of the curved sub-cell path (P1 and P2 on Q2 quads,
// SSP-RK2 and RK3).  The tapes of step n, stage s: WENO's limited and input
// DOFs and the corner impedances' |du.n| [(8 nd + 32) nc doubles], the
// clamp branches and the limited flags [44 nc bytes].  A segment slot keeps
// per stage the limited DOFs, the stage coordinates, u*, the corner records
// and the face slots.

static bool adjc_supported(const lag2d_ctx *x)
{
    const lag2d_config &c = x->cfg;
    return c.ncurv == 1 && c.nfint == 4 && (c.npoly == 1 || c.npoly == 2) && c.nrz == 0 &&
           (c.nstag == 2 || c.nstag == 3) && x->ntria == 0 && x->nrun == 1;
}

static size_t adjc_tape_doubles(const lag2d_ctx *x) { return (size_t)(8 * x->nd + 32) * x->nc; }
static size_t adjc_tape_bytes(const lag2d_ctx *x) { return (size_t)44 * x->nc; }
// corner states of the curved path: Q2_REC per cell, in tiles of 64 cells
static size_t adjc_rec_doubles(const lag2d_ctx *x) { return (size_t)Q2_REC * ((x->nc + 63) / 64 * 64); }
static size_t adjc_stage_stride(const lag2d_ctx *x)
{
    return (size_t)(4 * x->nd + 48) * x->nc + adjc_rec_doubles(x) + (size_t)4 * x->np;
}

// One step of the curved path's recorded (RECORD), replayed (REPLAY) or
// recomputed (SEGMENT, keeping the stage data in segment slot k) forward.
static int adjc_fwd_step(lag2d_ctx *x, int itime, int n, int mode, int k)
{
    Dev &d = x->d;
    lag2d_adj &A = x->adj;
    const int nc = x->nc, np = x->np, ns = x->cfg.nstag;
    const size_t nU = (size_t)4 * x->nd * nc, sst = adjc_stage_stride(x);
    auto cp = [&](double *dst, const double *src, size_t m) {
        return hipMemcpyAsync(dst, src, m * sizeof(double), hipMemcpyDeviceToDevice, x->stream);
    };
    for (int s = 0; s < ns; ++s) {
        const double alfa = RKLAG[ns - 2][s];
        Dev ds = d;
        if (s == 0) {
            ds.xs = d.xn;
            ds.ys = d.yn;
        }
        const double *Uin = (s == 0) ? d.Un : d.Us;
        const size_t ts = (size_t)n * ns + s;
        ds.wdu = A.lim_tape + ts * adjc_tape_doubles(x);
        ds.dlt = ds.wdu + 2 * nU;
        ds.ccl = A.clamp_tape + ts * adjc_tape_bytes(x);
        ds.dcl = ds.ccl + (size_t)18 * nc;
        double *slot = A.seg + (size_t)k * A.seg_stride + (size_t)s * sst;
        if (mode == ADJ_SEGMENT) {
            HIPCHK(cp(slot + nU, ds.xs, np));
            HIPCHK(cp(slot + nU + np, ds.ys, np));
        }
        int rc = launch_rhs_q2_adj(x, ds, Uin, alfa, s > 0, mode == ADJ_RECORD);
        if (rc) return fail(rc, "curved adjoint forward");
        if (mode == ADJ_SEGMENT) {
            HIPCHK(cp(slot, d.Ul, nU));
            HIPCHK(cp(slot + nU + 2 * (size_t)np, d.usx, np));
            HIPCHK(cp(slot + nU + 3 * (size_t)np, d.usy, np));
            HIPCHK(cp(slot + nU + 4 * (size_t)np, d.rec, adjc_rec_doubles(x)));
            HIPCHK(cp(slot + nU + 4 * (size_t)np + adjc_rec_doubles(x), d.gq, (size_t)48 * nc));
        }
        if (s == 0) {
            if (mode == ADJ_RECORD) {
                LAUNCH(KID_DT_PARTIAL, k_dt_partial, x->nblk_dt, ds, Uin, x->nblk_dt);
                LAUNCH(KID_DT_FINAL, k_dt_final, 1, d, itime, x->nblk_dt);
                HIPCHK(cp(A.dt_tape + n, d.scal, 1));
            } else {
                LAUNCH(KID_DT_FINAL, k_replay_dt, 1, d, A.dt_tape + n);
            }
            if (x->update_cells)
                LAUNCH(KID_UPDATE, k_update_cells, nblocks(4 * nc), d, alfa);
            else
                LAUNCH(KID_UPDATE, k_update, nblocks(4 * x->nd * nc), d, alfa);
        }
        LAUNCH(KID_NODEMOVE, k_nodemove, nblocks(np), d, alfa, s == 0 ? d.xn : d.xs, s == 0 ? d.yn : d.ys);
        launch_cellvol(x, d.xs, d.ys, 1, true);
    }
    HIPCHK(hipGetLastError());
    std::swap(d.Un, d.Us);
    std::swap(d.xn, d.xs);
    std::swap(d.yn, d.ys);
    x->parity ^= 1;
    return 0;
}

static int adjc_record(lag2d_ctx *x, int itime0, int nsteps, int every)
{
    const int nc = x->nc, np = x->np, ns = x->cfg.nstag;
    lag2d_adj &A = x->adj;
    const int nckpt = (nsteps + every - 1) / every;
    const size_t nU = (size_t)4 * x->nd * nc, ckw = nU + 4 * (size_t)np;
    int rc;
    if (!A.curved || A.cap_steps < (size_t)nsteps) {
        if ((rc = x->alloc(&A.lim_tape, (size_t)nsteps * ns * adjc_tape_doubles(x))) ||
            (rc = x->alloc(&A.clamp_tape, (size_t)nsteps * ns * adjc_tape_bytes(x))) ||
            (rc = x->alloc(&A.dt_tape, (size_t)nsteps)))
            return rc;
        A.cap_steps = nsteps;
    }
    if (!A.curved || A.cap_ckpt < (size_t)nckpt) {
        if ((rc = x->alloc(&A.ck, (size_t)nckpt * ckw))) return rc;
        A.cap_ckpt = nckpt;
    }
    if (!A.curved || A.cap_seg < (size_t)every) {
        A.seg_stride = (size_t)ns * adjc_stage_stride(x);
        if ((rc = x->alloc(&A.seg, (size_t)every * A.seg_stride)) ||
            (rc = x->alloc(&A.buf, 3 * nU + (size_t)11 * np + (size_t)(32 + 18 + 160 + 96) * nc)))
            return rc;
        A.cap_seg = every;
    }
    A.curved = true;
    A.nsteps = nsteps;
    A.every = every;
    A.nckpt = nckpt;
    A.itime0 = itime0;
    Dev &d = x->d;
    for (int n = 0; n < nsteps; ++n) {
        if (n % every == 0) {
            double *c = A.ck + (size_t)(n / every) * ckw;
            HIPCHK(hipMemcpyAsync(c, d.Un, nU * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
            HIPCHK(hipMemcpyAsync(c + nU, d.xn, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
            HIPCHK(hipMemcpyAsync(c + nU + np, d.yn, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
            HIPCHK(hipMemcpyAsync(c + nU + 2 * (size_t)np, d.usx, np * sizeof(double), hipMemcpyDeviceToDevice,
                                  x->stream));
            HIPCHK(hipMemcpyAsync(c + nU + 3 * (size_t)np, d.usy, np * sizeof(double), hipMemcpyDeviceToDevice,
                                  x->stream));
        }
        rc = adjc_fwd_step(x, itime0 + n, n, ADJ_RECORD, 0);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

static int adjc_gradient(lag2d_ctx *x, const double *lam_coord, double *grad)
{
    lag2d_adj &A = x->adj;
    Dev &d = x->d;
    const int nc = x->nc, np = x->np, ns = x->cfg.nstag, nd = x->nd;
    const size_t nU = (size_t)4 * nd * nc, ckw = nU + 4 * (size_t)np, sst = adjc_stage_stride(x);
    AdjStageC a{};
    double *b = A.buf;
    a.LU = b; b += nU;
    a.LUN = b; b += nU;
    a.LUI = b; b += nU;
    a.LX = b; b += 2 * (size_t)np;
    a.LXN = b; b += 2 * (size_t)np;
    a.LXI = b; b += 2 * (size_t)np;
    a.LUS = b; b += 2 * (size_t)np;
    a.NG = b; b += 3 * (size_t)np;
    a.CU = b; b += 32 * (size_t)nc;
    a.CX = b; b += 18 * (size_t)nc;
    a.LREC = b; b += 160 * (size_t)nc;
    a.LUSE = b; b += 96 * (size_t)nc;
    HIPCHK(hipMemsetAsync(A.buf, 0, (size_t)(b - A.buf) * sizeof(double), x->stream));
    std::vector<double> h(2 * (size_t)np);
    for (int p = 0; p < np; ++p) {
        h[p] = lam_coord[2 * (size_t)p];
        h[np + p] = lam_coord[2 * (size_t)p + 1];
    }
    HIPCHK(hipMemcpyAsync(a.LX, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
    for (int sg = A.nckpt - 1; sg >= 0; --sg) {
        const int n0 = sg * A.every, n1 = std::min(n0 + A.every, A.nsteps);
        const double *c = A.ck + (size_t)sg * ckw;
        HIPCHK(hipMemcpyAsync(d.Un, c, nU * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.xn, c + nU, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.yn, c + nU + np, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.usx, c + nU + 2 * (size_t)np, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.usy, c + nU + 3 * (size_t)np, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        for (int n = n0; n < n1; ++n) {
            int rc = adjc_fwd_step(x, A.itime0 + n, n, ADJ_SEGMENT, n - n0);
            if (rc) return rc;
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const double *slot0 = A.seg + (size_t)(n - n0) * A.seg_stride;
            a.dt = A.dt_tape + n;
            for (int s = ns - 1; s >= 0; --s) {
                const double *slot = slot0 + (size_t)s * sst;
                const size_t ts = (size_t)n * ns + s;
                a.Ul = slot;
                a.xin = slot + nU;
                a.yin = a.xin + np;
                a.usx = a.yin + np;
                a.usy = a.usx + np;
                a.rec = a.usy + np;
                a.gq = a.rec + adjc_rec_doubles(x);
                a.dlt = A.lim_tape + ts * adjc_tape_doubles(x) + 2 * nU;
                a.ccl = A.clamp_tape + ts * adjc_tape_bytes(x);
                a.dcl = a.ccl + (size_t)18 * nc;
                int rc = launch_adjq2_stage(x, a, RKLAG[ns - 2][s]);
                if (rc) return fail(rc, "curved adjoint stage");
                if (s > 0) {
                    std::swap(a.LU, a.LUI);
                    std::swap(a.LX, a.LXI);
                }
            }
            int rc = launch_adjq2_step_end(x, a);
            if (rc) return fail(rc, "curved adjoint step end");
        }
    }
    A.lam_x0 = a.LX;
    std::vector<double> g(nU);
    HIPCHK(hipMemcpyAsync(g.data(), a.LU, nU * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int ie = 0; ie < nc; ++ie)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < nd; ++id)
                grad[id + nd * (q + 4 * (size_t)ie)] = g[((size_t)q * nd + id) * nc + ie];
    return 0;
}

extern "C" int lag2d_adj_record(lag2d_ctx *x, int itime0, int nsteps, int every)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    if (nsteps < 1 || every < 1 || itime0 < 1) return fail(LAG2D_EINVAL, "nsteps, every and itime0 must be >= 1");
    if (adjc_supported(x)) return adjc_record(x, itime0, nsteps, every);
    if (!adj_supported(x))
        return fail(LAG2D_EUNSUPPORTED, "the adjoint covers the straight P1 planar path on quads with SSP-RK2 "
                                        "and the curved sub-cell path (P1, P2) on Q2 quads with SSP-RK2/3");
    const int nc = x->nc, np = x->np;
    lag2d_adj &A = x->adj;
    const int nckpt = (nsteps + every - 1) / every;
    if (A.nsteps < nsteps || A.every < every || A.nckpt < nckpt) {
        int rc;
        if ((rc = x->alloc(&A.lim_tape, (size_t)nsteps * 34 * nc)) || (rc = x->alloc(&A.dt_tape, (size_t)nsteps)) ||
            (rc = x->alloc(&A.ck, (size_t)nckpt * (12 * (size_t)nc + 4 * (size_t)np))))
            return rc;
        A.seg_stride = seg_stride(x);
        if ((rc = x->alloc(&A.seg, (size_t)every * A.seg_stride)) ||
            (rc = x->alloc(&A.buf, (size_t)(36 + 8 * 2 + 40) * nc + (size_t)13 * np)))
            return rc;
    }
    A.nsteps = nsteps;
    A.every = every;
    A.nckpt = nckpt;
    A.itime0 = itime0;
    Dev &d = x->d;
    const size_t ckw = 12 * (size_t)nc + 4 * (size_t)np;
    for (int n = 0; n < nsteps; ++n) {
        if (n % every == 0) {
            double *c = A.ck + (size_t)(n / every) * ckw;
            HIPCHK(hipMemcpyAsync(c, d.Un, 12 * (size_t)nc * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
            HIPCHK(hipMemcpyAsync(c + 12 * (size_t)nc, d.xn, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
            HIPCHK(hipMemcpyAsync(c + 12 * (size_t)nc + np, d.yn, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
            HIPCHK(hipMemcpyAsync(c + 12 * (size_t)nc + 2 * (size_t)np, d.usx, np * sizeof(double), hipMemcpyDeviceToDevice,
                                  x->stream));
            HIPCHK(hipMemcpyAsync(c + 12 * (size_t)nc + 3 * (size_t)np, d.usy, np * sizeof(double), hipMemcpyDeviceToDevice,
                                  x->stream));
        }
        int rc = adj_fwd_step(x, itime0 + n, n, ADJ_RECORD, 0);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

extern "C" int lag2d_adj_replay(lag2d_ctx *x, int nsteps)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    if (x->adj.nsteps < 1) return fail(LAG2D_EINVAL, "nothing recorded (lag2d_adj_record)");
    if (nsteps < 0 || nsteps > x->adj.nsteps) return fail(LAG2D_EINVAL, "more steps than recorded");
    for (int n = 0; n < nsteps; ++n) {
        int rc = x->adj.curved ? adjc_fwd_step(x, x->adj.itime0 + n, n, ADJ_REPLAY, 0)
                               : adj_fwd_step(x, x->adj.itime0 + n, n, ADJ_REPLAY, 0);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

// One step reversed from the stage data of segment slot k.
static int adj_rev_step(lag2d_ctx *x, int n, int k, AdjStage &a)
{
    Dev &d = x->d;
    lag2d_adj &A = x->adj;
    const int nc = x->nc, np = x->np;
    const size_t nU = (size_t)12 * nc;
    const double *slot = A.seg + (size_t)k * A.seg_stride;
    const double *Xn = slot + 2 * nU, *X1 = Xn + 2 * (size_t)np, *usp = X1 + 2 * (size_t)np;
    const double *us0 = usp + 2 * (size_t)np, *us1 = us0 + 2 * (size_t)np;
    const double *rec0 = us1 + 2 * (size_t)np, *rec1 = rec0 + 40 * (size_t)nc;
    a.dt = A.dt_tape + n;
    for (int s = 1; s >= 0; --s) {
        const double alfa = RKLAG[0][s];
        a.Uin = s == 1 ? slot + nU : slot;
        a.xin = s == 1 ? X1 : Xn;
        a.yin = a.xin + np;
        a.usx = s == 1 ? us1 : us0;
        a.usy = a.usx + np;
        a.rec = s == 1 ? rec1 : rec0;
        a.lim = A.lim_tape + ((size_t)n * 2 + s) * 17 * nc;
        a.dlt = a.lim + 9 * (size_t)nc;
        LAUNCH(KID_ADJ_NODES, k_adj_nodemove, nblocks(np), d, a, alfa);
        LAUNCH(KID_ADJ_CELLRHS, k_adj_cellrhs, nblocks(nc), d, a, alfa);
        LAUNCH(KID_ADJ_NODES, k_adj_node, nblocks(np), d, a);
        LAUNCH(KID_ADJ_CORNER, k_adj_corner, nblocks(nc), d, a);
        LAUNCH(KID_ADJ_NODES, k_adj_gather, nblocks(np), d, a);
        if (s == 1) std::swap(a.LU, a.LUI);
    }
    LAUNCH(KID_ADJ_NODES, k_adj_step_end, nblocks(std::max(12 * nc, 2 * np)), d, a);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" int lag2d_adj_gradient(lag2d_ctx *x, const double *lam_coord, double *grad)
{
    if (!x || !lam_coord || !grad) return fail(LAG2D_EINVAL, "null argument");
    lag2d_adj &A = x->adj;
    if (A.nsteps < 1) return fail(LAG2D_EINVAL, "nothing recorded (lag2d_adj_record)");
    if (A.curved) return adjc_gradient(x, lam_coord, grad);
    Dev &d = x->d;
    const int nc = x->nc, np = x->np;
    const size_t nU = (size_t)12 * nc;
    AdjStage a{};
    double *b = A.buf;
    a.LU = b; b += nU;
    a.LUN = b; b += nU;
    a.LUI = b; b += nU;
    a.CX = b; b += 8 * (size_t)nc;
    a.CU = b; b += 8 * (size_t)nc;
    a.LREC = b; b += 40 * (size_t)nc;
    a.LX = b; b += 2 * (size_t)np;
    a.LXN = b; b += 2 * (size_t)np;
    a.LXI = b; b += 2 * (size_t)np;
    a.LUS = b; b += 2 * (size_t)np;
    a.LUSN = b; b += 2 * (size_t)np;
    a.NG = b; b += 3 * (size_t)np;
    HIPCHK(hipMemsetAsync(A.buf, 0, (size_t)(b - A.buf) * sizeof(double), x->stream));
    std::vector<double> h(2 * (size_t)np);
    for (int p = 0; p < np; ++p) {
        h[p] = lam_coord[2 * (size_t)p];
        h[np + p] = lam_coord[2 * (size_t)p + 1];
    }
    HIPCHK(hipMemcpyAsync(a.LX, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, x->stream));
    const size_t ckw = 12 * (size_t)nc + 4 * (size_t)np;
    for (int sg = A.nckpt - 1; sg >= 0; --sg) {
        const int n0 = sg * A.every, n1 = std::min(n0 + A.every, A.nsteps);
        const double *c = A.ck + (size_t)sg * ckw;
        HIPCHK(hipMemcpyAsync(d.Un, c, nU * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.xn, c + nU, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.yn, c + nU + np, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.usx, c + nU + 2 * (size_t)np, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        HIPCHK(hipMemcpyAsync(d.usy, c + nU + 3 * (size_t)np, np * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
        for (int n = n0; n < n1; ++n) {
            int rc = adj_fwd_step(x, A.itime0 + n, n, ADJ_SEGMENT, n - n0);
            if (rc) return rc;
        }
        for (int n = n1 - 1; n >= n0; --n) {
            int rc = adj_rev_step(x, n, n - n0, a);
            if (rc) return rc;
        }
    }
    A.lam_x0 = a.LX;
    std::vector<double> g(nU);
    HIPCHK(hipMemcpyAsync(g.data(), a.LU, nU * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int ie = 0; ie < nc; ++ie)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) grad[id + 3 * (q + 4 * (size_t)ie)] = g[((size_t)q * 3 + id) * nc + ie];
    return 0;
}

extern "C" int lag2d_adj_gradient_coord(lag2d_ctx *x, double *grad_coord)
{
    if (!x || !grad_coord) return fail(LAG2D_EINVAL, "null argument");
    if (!x->adj.lam_x0) return fail(LAG2D_EINVAL, "no gradient computed (lag2d_adj_gradient)");
    const int np = x->np;
    std::vector<double> h(2 * (size_t)np);
    HIPCHK(hipMemcpyAsync(h.data(), x->adj.lam_x0, h.size() * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    for (int p = 0; p < np; ++p) {
        grad_coord[2 * (size_t)p] = h[p];
        grad_coord[2 * (size_t)p + 1] = h[np + p];
    }
    return 0;
}

// A step through a captured HIP graph: the first step of each key (swap
// parity, fixed dt and epsaw) is captured from step_impl's launches, later
// ones replay it.  The first step of a run (dt from the CFL alone), timed
// steps and curved-path steps before the variants are tuned run eagerly.
static int step_graph(lag2d_ctx *x, int itime)
{
    const bool tuned = x->cfg.ncurv != 1 || (x->corner_variant != 0 && x->rhs_variant != 0);
    if (!x->use_graphs || itime <= 1 || x->timing || !tuned) return step_impl(x, itime);
    double dt = 0.0, epsaw = 0.0;
    const int sched = schedule(x->cfg, itime, &dt, &epsaw) ? 1 : 0;
    Dev &d = x->d;
    for (auto &g : x->graphs) {
        if (g.parity == x->parity && g.sched == sched && g.dt == dt && g.epsaw == epsaw) {
            HIPCHK(hipGraphLaunch(g.exec, x->stream));
            std::swap(d.Un, d.Us);
            std::swap(d.xn, d.xs);
            std::swap(d.yn, d.ys);
            x->parity ^= 1;
            return 0;
        }
    }
    lag2d_graph g;
    g.parity = x->parity;
    g.sched = sched;
    g.dt = dt;
    g.epsaw = epsaw;
    HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
    const int rc = step_impl(x, itime);  // records the launches and swaps the buffers
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(x->stream, &graph);
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    HIPCHK(e);
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    HIPCHK(ei);
    x->graphs.push_back(g);
    HIPCHK(hipGraphLaunch(g.exec, x->stream));
    return 0;
}

extern "C" int lag2d_use_graphs(lag2d_ctx *x, int enable)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    x->use_graphs = enable != 0;
    return 0;
}

// Strict mode on the curved path: the reference stops where its WENO stencil
// builder meets a vertex neighbourhood it cannot classify (WENO.f90:7699);
// the kernels flag such troubled cells in wenobad, read back here (this
// synchronises the stream, so strict steps are not asynchronous).
static int check_strict(lag2d_ctx *x, int itime)
{
    if (!(x->cfg.strict && x->cfg.ncurv == 1)) return 0;
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, x->d.wenobad, sizeof(int), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (bad)
        return fail(LAG2D_ETOPOLOGY, "WENO stencil: troubled cell with a vertex neighbourhood the "
                                     "reference stops on (WENO.f90:7699) at step " +
                                         std::to_string(itime));
    return 0;
}

extern "C" int lag2d_step(lag2d_ctx *x, int itime)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    if (itime < 1) return fail(LAG2D_EINVAL, "itime is 1-based");
    const int rc = step_graph(x, itime);
    if (rc) return rc;
    return check_strict(x, itime);
}

extern "C" int lag2d_run(lag2d_ctx *x, int itime0, int nsteps, int *taken)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    int n = 0;
    // With a CFL-controlled dt the end-time test reads tacru back after every
    // step (the driver stops at the step that reaches tend); lag2d_step queues
    // steps without any host synchronisation.
    double sc[2];
    HIPCHK(hipMemcpyAsync(sc, x->d.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    bool watch = x->cfg.ndt == 1;
    for (int k = 0; k < nsteps; ++k) {
        int rc = step_graph(x, itime0 + k);
        if (rc) return rc;
        ++n;
        if ((rc = check_strict(x, itime0 + k))) return rc;
        if (watch && x->nrun == 1) {
            HIPCHK(hipMemcpyAsync(sc, x->d.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, x->stream));
            HIPCHK(hipStreamSynchronize(x->stream));
            if (std::fabs(sc[1] - x->cfg.tend) / x->cfg.tend <= 1e-11) break;
        } else if (x->nrun > 1) {
            // an ensemble ends when every run has reached its end time
            std::vector<double> all((size_t)2 * x->nrun);
            HIPCHK(hipMemcpyAsync(all.data(), x->d.scal, all.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  x->stream));
            HIPCHK(hipStreamSynchronize(x->stream));
            bool done = true;
            for (int r = 0; r < x->nrun && done; ++r)
                done = std::fabs(all[2 * r + 1] - x->tend_host[r]) / x->tend_host[r] <= 1e-11;
            if (done) break;
        }
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    if (taken) *taken = n;
    return 0;
}

extern "C" int lag2d_rhs(lag2d_ctx *x, double *rhsel)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    Dev &d = x->d;
    Dev ds = d;
    ds.xs = d.xn;
    ds.ys = d.yn;
    int rc = launch_rhs(x, ds, d.Un, 0.0, 0, 1, 0);
    if (rc) return rc;
    if ((rc = check_strict(x, 1))) return rc;
    const int nc = x->nc, nd = x->nd;
    const size_t nU = (size_t)4 * nd * nc;
    HIPCHK(hipMemcpyAsync(d.Un, d.Us, nU * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (rhsel) {
        std::vector<double> h(nU);
        HIPCHK(hipMemcpy(h.data(), d.R, nU * sizeof(double), hipMemcpyDeviceToHost));
        for (int ie = 0; ie < nc; ++ie)
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < nd; ++id)
                    rhsel[id + (size_t)nd * (q + 4 * (size_t)ie)] = h[((size_t)q * nd + id) * nc + ie];
    }
    return 0;
}

extern "C" int lag2d_download(lag2d_ctx *x, double *coord, double *geoel, double *unkno, double *ustar)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(x->stream));
    const int nc = x->nc, np = x->np, nd = x->nd;
    Dev &d = x->d;
    std::vector<double> h((size_t)std::max({2 * np, 4 * nd * nc, 24 * nc}));
    if (coord) {
        HIPCHK(hipMemcpy(h.data(), d.xn, np * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h.data() + np, d.yn, np * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < np; ++p) {
            coord[2 * (size_t)p] = h[p];
            coord[2 * (size_t)p + 1] = h[np + p];
        }
    }
    if (ustar) {
        HIPCHK(hipMemcpy(h.data(), d.usx, np * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(h.data() + np, d.usy, np * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < np; ++p) {
            ustar[2 * (size_t)p] = h[p];
            ustar[2 * (size_t)p + 1] = h[np + p];
        }
    }
    if (unkno) {
        HIPCHK(hipMemcpy(h.data(), d.Un, (size_t)4 * nd * nc * sizeof(double), hipMemcpyDeviceToHost));
        for (int ie = 0; ie < nc; ++ie)
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < nd; ++id)
                    unkno[id + (size_t)nd * (q + 4 * (size_t)ie)] = h[((size_t)q * nd + id) * nc + ie];
    }
    if (geoel) {
        HIPCHK(hipMemcpy(h.data(), d.G, (size_t)24 * nc * sizeof(double), hipMemcpyDeviceToHost));
        for (int ie = 0; ie < nc; ++ie)
            for (int k = 0; k < 24; ++k) {
                bool live = (k == 2 || k == 4 || k == 5) ||
                            (x->cfg.nrz == 2 && (k == 0 || k == 1 || (k >= 10 && k <= 15)));
                geoel[k + 24 * (size_t)ie] = live ? h[(size_t)k * nc + ie] : x->geo_host[k + 24 * (size_t)ie];
            }
    }
    return 0;
}

extern "C" int lag2d_get_mass_matrix(lag2d_ctx *x, double *amatr)
{
    if (!x || !amatr) return fail(LAG2D_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(x->stream));
    const int nc = x->nc, nm = x->nmatr;
    std::vector<double> h((size_t)nm * nc);
    HIPCHK(hipMemcpy(h.data(), x->A, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int ie = 0; ie < nc; ++ie)
        for (int k = 0; k < nm; ++k) amatr[k + (size_t)nm * ie] = h[(size_t)k * nc + ie];
    return 0;
}

extern "C" int lag2d_set_initial_density(lag2d_ctx *x, const double *rho)
{
    if (!x || !rho) return fail(LAG2D_EINVAL, "null argument");
    if (x->cfg.ncurv != 1) return fail(LAG2D_EUNSUPPORTED, "the initial density is used by the curved P2 path");
    Dev &d = x->d;
    if (!d.rhoi) {
        double *t;
        int rc = x->alloc(&t, (size_t)x->nc);
        if (rc) return rc;
        d.rhoi = t;
    }
    HIPCHK(hipMemcpy((void *)d.rhoi, rho, (size_t)x->nc * sizeof(double), hipMemcpyHostToDevice));
    int rc = upload_extra(*x);  // initial sub-cell masses from this density
    if (rc) return fail(rc, "initial sub-cell masses");
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

extern "C" int lag2d_weno_cells(lag2d_ctx *x, long long *total)
{
    if (!x || !total) return fail(LAG2D_EINVAL, "null argument");
    *total = 0;
    if (!x->d.tsum) return 0;
    unsigned long long v = 0;
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(&v, x->d.tsum, sizeof(v), hipMemcpyDeviceToHost));
    *total = (long long)v;
    return 0;
}

extern "C" int lag2d_division_check(int n, const double *a, const double *b, double *q)
{
    if (n < 0 || (n > 0 && (!a || !b || !q))) return fail(LAG2D_EINVAL, "bad argument");
    if (n == 0) return 0;
    double *da = nullptr;
    int rc = dalloc(&da, (size_t)3 * n);
    if (rc) return rc;
    const size_t bytes = (size_t)n * sizeof(double);
    hipError_t e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(da + n, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_division_check, dim3(nblocks(n)), dim3(LAG2D_BLOCK), 0, 0, da, da + n, da + 2 * (size_t)n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(q, da + 2 * (size_t)n, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(da);
    if (e != hipSuccess) return fail(LAG2D_EDEVICE, std::string("division check: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int lag2d_get_run_times(lag2d_ctx *x, double *dtfix, double *tacru)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    std::vector<double> sc((size_t)2 * x->nrun);
    HIPCHK(hipMemcpy(sc.data(), x->d.scal, sc.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < x->nrun; ++r) {
        if (dtfix) dtfix[r] = sc[2 * r];
        if (tacru) tacru[r] = sc[2 * r + 1];
    }
    return 0;
}

extern "C" int lag2d_get_time(lag2d_ctx *x, double *dtfix, double *tacru)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    double sc[2];
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(sc, x->d.scal, sizeof(sc), hipMemcpyDeviceToHost));
    if (dtfix) *dtfix = sc[0];
    if (tacru) *tacru = sc[1];
    return 0;
}

extern "C" int lag2d_check(lag2d_ctx *x)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    const int nc = x->nc, nd = x->nd;
    std::vector<double> u((size_t)4 * nd * nc), g((size_t)3 * nc);
    HIPCHK(hipStreamSynchronize(x->stream));
    HIPCHK(hipMemcpy(u.data(), x->d.Un, u.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(g.data(), x->d.G + 2 * (size_t)nc, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < u.size(); ++k)
        if (!std::isfinite(u[k]))
            return fail(LAG2D_ENONFINITE, "non-finite DOF in the state at cell " + std::to_string(k % nc + 1));
    for (int ie = 0; ie < nc; ++ie)
        if (!(g[ie] > 0.0)) return fail(LAG2D_EVOLUME, "non-positive cell volume at cell " + std::to_string(ie + 1));
    // the cell average of tau (specific volume, variable 1, degree 1)
    for (int ie = 0; ie < nc; ++ie)
        if (!(u[ie] > 0.0))
            return fail(LAG2D_EVOLUME, "non-positive specific volume (mean tau) at cell " + std::to_string(ie + 1));
    return 0;
}

extern "C" int lag2d_sync(lag2d_ctx *x)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(x->stream));
    return 0;
}

extern "C" int lag2d_kernel_time(lag2d_ctx *x, int enable, double *avg_ms, const char **name)
{
    if (!x) return fail(LAG2D_EINVAL, "null context");
    if (enable >= 0) {
        HIPCHK(hipStreamSynchronize(x->stream));
        for (auto &t : x->timers) {
            for (auto &ev : t.ev) {
                hipEventDestroy(ev.first);
                hipEventDestroy(ev.second);
            }
            t.ev.clear();
        }
        x->timing = enable != 0;
        if (enable && x->d.tsum) HIPCHK(hipMemset(x->d.tsum, 0, sizeof(unsigned long long)));
        return 0;
    }
    // enable < 0: report
    HIPCHK(hipStreamSynchronize(x->stream));
    int best = -1;
    double best_tot = 0.0, best_avg = 0.0;
    for (int k = 0; k < KID_COUNT; ++k) {
        auto &t = x->timers[k];
        if (t.ev.empty()) continue;
        double tot = 0.0;
        for (auto &ev : t.ev) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev.first, ev.second);
            tot += ms;
        }
        x->kernel_total_ms[k] = tot;
        x->kernel_calls[k] = (int)t.ev.size();
        if (tot > best_tot) {
            best_tot = tot;
            best_avg = tot / t.ev.size();
            best = k;
        }
    }
    if (best < 0) return fail(LAG2D_EINVAL, "no timed launches");
    if (avg_ms) *avg_ms = best_avg;
    if (name) *name = kernel_name(x, best);
    return 0;
}

extern "C" int lag2d_kernel_stats(lag2d_ctx *x, int k, double *total_ms, int *calls, const char **name)
{
    if (!x || k < 0 || k >= KID_COUNT) return fail(LAG2D_EINVAL, "bad argument");
    if (total_ms) *total_ms = x->kernel_total_ms[k];
    if (calls) *calls = x->kernel_calls[k];
    if (name) *name = kernel_name(x, k);
    return 0;
}

extern "C" void lag2d_destroy(lag2d_ctx *x) { delete x; }
