// lee_kernels.hip -- HIP kernels of the FB_p_v time step for gfx950 (MI355X).
//
// One thread per grid point, x (i) across the 64 lanes of a wavefront so every field
// access of a wavefront is one contiguous 512 B (FP64) / 256 B (FP32) segment; the
// y/z neighbours are re-read from L2 (the radius-1 stencil reuses each line three times
// per plane).  Built with -ffp-contract=off: the FP64 path keeps the reference's
// association order (air.cpp / porous.cpp), so it reproduces the CPU results bit for bit.
//
// The reference grid is separable (cross derivatives X_over_y ... are fixed at zero,
// calculation_2D.cpp:102-104): of the nine inverse-metric terms only x_over_X, y_over_Y,
// z_over_Z are non-zero, and every product with a zero term is dropped here.  Dropping
// an exact zero from a sum leaves the value unchanged (only the sign of an exact-zero
// result can differ).
#include "lee_kernels.h"

#include <algorithm>
#include <climits>
#include <utility>
#include <vector>
#include <type_traits>

namespace lee {

namespace {

constexpr int BX = 64, BY = 4;
static_assert(BX * BY == kBatchBlock, "ensemble blocks are BX x BY threads");

// Tile t of a flat grid whose blocks go round-robin to the 8 XCDs: XCD x runs the x-th
// contiguous range of tiles, in order.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int x = b & 7, per = nb >> 3, rem = nb & 7;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + (b >> 3);
}


// A box of 64 x 4-point tiles: tile columns bx .. bx+nbx-1 (points i = 64 bx ...), tile
// rows by .. by+nby-1 (points j = 4 by + 1 ...), planes k .. k+nk-1.  The split launches
// cover the tile space of a domain with lists of such boxes, one list per set of slab axes
// its points can meet (see split_regions).
struct TileBox {
    int bx, nbx, by, nby, k, nk;
};
constexpr int kMaxBoxes = 18;
template <int NB> struct BoxList {
    TileBox b[NB];
    int first[NB + 1];  // first block of each box; first[n] = blocks in all, then INT_MAX
    int n;
    int pad;  // each plane of a box takes a multiple of 8 blocks (list_point)
};

// Point of block t of a list: thread (threadIdx.x, threadIdx.y) of the box's tiles, taken x
// fastest, then y, then planes.  blockIdx is uniform, so this is scalar arithmetic on the
// argument, read at constant offsets only: the box's fields as sums of products with 0/1
// (selects between loaded values were merged into one load through a selected address,
// which put the argument in private memory).  With pad each plane of a box takes a multiple
// of 8 blocks (the extra ones get i = -1 and leave at the bounds check) and the boxes start at
// multiples of 8: blocks are dealt to the 8 XCDs round-robin, so a tile column then stays on
// one XCD through its planes, and the planes k - 1, k + 1 a tile reads come from that XCD's
// L2 (on the 318-tile planes of the 512 x 256 x 128 interior, plane k + 1 of a column went to
// another XCD and the RK4 stage fetched twice its compulsory reads).  It takes a third off
// the RK4 stage's HBM reads at the same speed; the FB passes ran 2 % slower with it and
// stay unpadded.
template <int NB, int... Q>
__device__ __forceinline__ void list_point_(const BoxList<NB> &l, int t, int &i, int &j, int &k,
                                            std::integer_sequence<int, Q...>)
{
    const int q = (0 + ... + (int)(Q > 0 && t >= l.first[Q]));
    const int bx = (0 + ... + ((q == Q) * l.b[Q].bx)), nbx = (0 + ... + ((q == Q) * l.b[Q].nbx));
    const int by = (0 + ... + ((q == Q) * l.b[Q].by)), nby = (0 + ... + ((q == Q) * l.b[Q].nby));
    const int k0 = (0 + ... + ((q == Q) * l.b[Q].k)), f = (0 + ... + ((q == Q) * l.first[Q]));
    t -= f;
    const int per = nbx * nby, pp = l.pad ? (per + 7) & ~7 : per, kk = t / pp, col = t - kk * pp, ry = col / nbx;
    if (col >= per) {  // padding: leaves at the caller's bounds check
        i = -1; j = 1; k = k0;
        return;
    }
    k = k0 + kk;
    i = (bx + (col - ry * nbx)) * 64 + threadIdx.x;
    j = (by + ry) * BY + threadIdx.y + 1;
}
template <int NB> __device__ __forceinline__ void list_point(const BoxList<NB> &l, int t, int &i, int &j, int &k)
{
    list_point_(l, t, i, j, k, std::make_integer_sequence<int, NB>{});
}

template <typename T> __device__ __forceinline__ void metric(const Grid<T> &g, int i, int j, int k, T &mX, T &mY, T &mZ)
{
    if constexpr (std::is_same<T, double>::value) {
        // Jacobian of calculation_2D.cpp:424-430 and Geo_over_var (df_over_var.cpp:5-9)
        double xx = g.mx[i], yy = g.my[j], zz = g.mz[k];
        double J = 1.0 / (xx * yy * zz);
        mX = (yy * zz) * J;
        mY = (xx * zz) * J;
        mZ = (xx * yy) * J;
    } else {
        mX = g.mx[i]; mY = g.my[j]; mZ = g.mz[k];
    }
}

// Terrain-following grid: X_over_x, Y_over_y as in the reference, Z_over_z = Zz[k]*sh(i,j),
// Z_over_x = hx(i,j)*tk[k], Z_over_y = hy(i,j)*tk[k].  Of the nine inverse-metric terms of
// Geo_over_var (df_over_var.cpp:5-9) five are non-zero: xX, yY, zZ and the cross terms
// zX = -(Yy*Zx)*J, zY = -(Xx*Zy)*J.  The tables hold X_over_x ... (not reciprocals) in both
// precisions.
template <typename T> __device__ __forceinline__ void metric_terrain(const Grid<T> &g, int i, int j, int k, T &mX, T &mY,
                                                                     T &mZ, T &zX, T &zY)
{
    const size_t c = (size_t)i + (size_t)g.nx * j;
    const T xx = g.mx[i], yy = g.my[j], zz = g.mz[k] * g.tsh[c];
    const T Zx = g.thx[c] * g.ttk[k], Zy = g.thy[c] * g.ttk[k];
    const T J = T(1) / (xx * yy * zz);
    mX = (yy * zz) * J;
    mY = (xx * zz) * J;
    mZ = (xx * yy) * J;
    zX = (-(yy * Zx)) * J;
    zY = (-(xx * Zy)) * J;
}

__device__ __forceinline__ size_t qxi(int ii, int j, int k, int n, int ny) { return (size_t)ii + (size_t)n * ((size_t)j + (size_t)ny * k); }
__device__ __forceinline__ size_t qyi(int i, int jj, int k, int nx, int n) { return (size_t)i + (size_t)nx * ((size_t)jj + (size_t)n * k); }
__device__ __forceinline__ size_t qzi(int i, int j, int kk, int nx, int ny) { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * kk); }

// Membership of interior point (i,j,k) in the six PML slabs and its index in each.
template <typename T> struct OwnQ {
    size_t idx[6];
    bool in[6];
};
// L as for dp_layers: slabs of the axes left out are known not to hold the point.
template <typename T, int L = 7> __device__ __forceinline__ OwnQ<T> own_slabs(const AirArgs<T> &a, int i, int j, int k)
{
    const int nx = a.g.nx, ny = a.g.ny;
    OwnQ<T> o;
    o.in[0] = (L & 1) && a.qn[0] && i < a.qlo[0] + a.qn[0];
    o.in[1] = (L & 1) && a.qn[1] && i >= a.qlo[1];
    o.in[2] = (L & 2) && a.qn[2] && a.y1on && j < a.qlo[2] + a.qn[2];
    o.in[3] = (L & 2) && a.qn[3] && j >= a.qlo[3];
    o.in[4] = (L & 4) && a.qn[4] && k < a.qlo[4] + a.qn[4];
    o.in[5] = (L & 4) && a.qn[5] && k >= a.qlo[5];
    o.idx[0] = o.in[0] ? qxi(i - a.qlo[0], j, k, a.qn[0], ny) : 0;
    o.idx[1] = o.in[1] ? qxi(i - a.qlo[1], j, k, a.qn[1], ny) : 0;
    o.idx[2] = o.in[2] ? qyi(i, j - a.qlo[2], k, nx, a.qn[2]) : 0;
    o.idx[3] = o.in[3] ? qyi(i, j - a.qlo[3], k, nx, a.qn[3]) : 0;
    o.idx[4] = o.in[4] ? qzi(i, j, k - a.qlo[4], nx, ny) : 0;
    o.idx[5] = o.in[5] ? qzi(i, j, k - a.qlo[5], nx, ny) : 0;
    return o;
}

// ------------------------------------------------------------------------------------
// The interior arithmetic of the pressure and velocity increments for the mean flows of
// velocity_method 0 and 7 (none, or V along y only), from the point's neighbour values.  The
// one-point stencils below and the four-point form of the RK4 stage both call these, so every
// launch shape computes the same bits.
// ------------------------------------------------------------------------------------
template <typename T> struct Nb {
    T mX, mY, mZ, zX, zY;                        // inverse metric (zX = zY = 0 on a flat grid)
    T pc, p_im, p_jp, p_jm, p_kp, p_km;          // p and its neighbours (i-1, j+-1, k+-1)
    T uc, u_ip, u_jp, u_jm, u_kp, u_km;          // u
    T vc, v_jp, v_jm, v_kp, v_km;                // v
    T wc, w_jp, w_jm, w_kp, w_km;                // w
    T Vc, V_ip, V_im, V_jp, V_jm, V_kp, V_km;    // mean flow along y
};

// dt*R_p without the layer terms (air.cpp:1953-2043); also the damped terms ddx, ddy and
// the difference dwz the layer terms use.
template <typename T, int WM, bool TERRAIN, bool CEN>
__device__ __forceinline__ T dp_core(const AirArgs<T> &a, const Nb<T> &n, T &ddx, T &ddy, T &dwz)
{
    const Grid<T> &g = a.g;
    const T dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, dt = g.dt;
    const T mX = n.mX, mY = n.mY, mZ = n.mZ, zX = n.zX, zY = n.zY, pc = n.pc;
    const T dux = n.u_ip - n.uc;
    const T dvy = n.v_jp - n.vc;
    dwz = n.w_kp - n.wc;
    T duz = T(0), dvz = T(0);
    T div;
    if constexpr (TERRAIN) {
        duz = n.u_kp - n.uc;
        dvz = n.v_kp - n.vc;
        div = (((((mX * dux) * dxi) + ((zX * duz) * dzi)) + ((mY * dvy) * dyi)) + ((zY * dvz) * dzi)) + ((mZ * dwz) * dzi);
    } else {
        div = (((mX * dux) * dxi) + ((mY * dvy) * dyi)) + ((mZ * dwz) * dzi);
    }
    ddx = (mX * dux) * dxi;
    ddy = (mY * dvy) * dyi;
    if constexpr (TERRAIN) {
        ddx = ddx + (zX * duz) * dzi;
        ddy = ddy + (zY * dvz) * dzi;
    }
    T r;
    if constexpr (WM == 2) {
        // mean flow along y only (U = W = 0): the terms with U or W are exact zeros
        const T Vc = n.Vc;
        const T c2 = ((Vc * mY) * dyi) * dt;
        T VY = ((mY * (n.V_jp - n.V_jm)) * dyi) * T(0.5);
        if constexpr (CEN) r = -(c2 * ((n.p_jp - n.p_jm) * T(0.5)));
        else r = -(c2 * (pc - n.p_jm));
        if constexpr (TERRAIN) {
            const T c3 = ((Vc * zY) * dzi) * dt;
            VY = VY + ((zY * (n.V_kp - n.V_km)) * dzi) * T(0.5);
            if constexpr (CEN) r = r - c3 * ((n.p_kp - n.p_km) * T(0.5));
            else r = r - c3 * (pc - n.p_km);
        }
        r += (((-a.gam) * pc) * dt) * VY;
        r += (((-a.gam) * a.Pav) * dt) * div;
    } else {
        r = (((-a.gam) * a.Pav) * dt) * div;
    }
    return r;
}

// dt*R_u, dt*R_v, dt*R_w without the layer terms (air.cpp:1476-1578); also the pressure
// gradients gpx, gpy, gpz the layer terms use.
template <typename T, int WM, bool TERRAIN, bool CEN>
__device__ __forceinline__ void duvw_core(const AirArgs<T> &a, const Nb<T> &n, T &gpx, T &gpy, T &gpz, T &du, T &dv,
                                          T &dw)
{
    const Grid<T> &g = a.g;
    const T dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, dt = g.dt, Aav = a.Aav;
    const T mX = n.mX, mY = n.mY, mZ = n.mZ, zX = n.zX, zY = n.zY, pc = n.pc;
    const T bpz = pc - n.p_km;
    gpx = (mX * (pc - n.p_im)) * dxi;
    gpy = (mY * (pc - n.p_jm)) * dyi;
    gpz = (mZ * bpz) * dzi;
    if constexpr (TERRAIN) {
        gpx = gpx + (zX * bpz) * dzi;
        gpy = gpy + (zY * bpz) * dzi;
    }
    const T uc = n.uc, vc = n.vc, wc = n.wc;
    if constexpr (WM == 2) {
        // mean flow along y only (U = W = 0): the terms with U or W or their gradients vanish
        const T Vc = n.Vc;
        const T c2 = ((Vc * mY) * dyi) * dt;
        T VX = ((mX * (n.V_ip - n.V_im)) * dxi) * T(0.5);
        T VY = ((mY * (n.V_jp - n.V_jm)) * dyi) * T(0.5);
        const T VZ = ((mZ * (n.V_kp - n.V_km)) * dzi) * T(0.5);
        // d/dY (and d/dZ over terrain) of one component: forward (reference) or central
        auto d1 = [&](T fp, T fm, T fc) {
            if constexpr (CEN) return (fp - fm) * T(0.5);
            else return fp - fc;
        };
        du = -(c2 * d1(n.u_jp, n.u_jm, uc));
        dv = -(c2 * d1(n.v_jp, n.v_jm, vc));
        dw = -(c2 * d1(n.w_jp, n.w_jm, wc));
        if constexpr (TERRAIN) {
            const T c3 = ((Vc * zY) * dzi) * dt;
            VX = VX + ((zX * (n.V_kp - n.V_km)) * dzi) * T(0.5);
            VY = VY + ((zY * (n.V_kp - n.V_km)) * dzi) * T(0.5);
            du = du - c3 * d1(n.u_kp, n.u_km, uc);
            dv = dv - c3 * d1(n.v_kp, n.v_km, vc);
            dw = dw - c3 * d1(n.w_kp, n.w_km, wc);
        }
        du += ((-Aav) * dt) * gpx;
        dv += (((-uc) * VX) * dt - (vc * VY) * dt) - (wc * VZ) * dt;
        dv += ((-Aav) * dt) * gpy;
        dw += ((-Aav) * dt) * gpz;
    } else {
        du = ((-Aav) * dt) * gpx;
        dv = ((-Aav) * dt) * gpy;
        dw = ((-Aav) * dt) * gpz;
    }
}

// Neighbour values of one point for dp_core (DP) and/or duvw_core (DUVW), read from memory;
// p of the pressure increment from g.pn, of the velocity increment from g.pnn.
template <typename T, int WM, bool TERRAIN, bool CEN, bool DP, bool DUVW>
__device__ __forceinline__ void load_nb(const AirArgs<T> &a, int i, int j, int k, Nb<T> &n)
{
    const Grid<T> &g = a.g;
    const size_t nx = (size_t)g.nx, sxy = (size_t)g.nx * g.ny;
    const size_t id = (size_t)i + nx * ((size_t)j + (size_t)g.ny * k);
    n.zX = T(0); n.zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, n.mX, n.mY, n.mZ, n.zX, n.zY);
    else metric(g, i, j, k, n.mX, n.mY, n.mZ);
    gptr<const T> u = g.un, v = g.vn, w = g.wn;
    n.uc = u[id]; n.vc = v[id]; n.wc = w[id];
    if constexpr (DP) {
        gptr<const T> p = g.pn;
        n.pc = p[id];
        n.u_ip = u[id + 1]; n.v_jp = v[id + nx]; n.w_kp = w[id + sxy];
        if constexpr (TERRAIN) { n.u_kp = u[id + sxy]; n.v_kp = v[id + sxy]; }
        if constexpr (WM == 2) {
            n.Vc = a.V[id]; n.V_jp = a.V[id + nx]; n.V_jm = a.V[id - nx];
            n.p_jm = p[id - nx];
            if constexpr (CEN) n.p_jp = p[id + nx];
            if constexpr (TERRAIN) {
                n.V_kp = a.V[id + sxy]; n.V_km = a.V[id - sxy];
                n.p_km = p[id - sxy];
                if constexpr (CEN) n.p_kp = p[id + sxy];
            }
        }
    }
    if constexpr (DUVW) {
        gptr<const T> p = g.pnn;
        n.pc = p[id];
        n.p_im = p[id - 1]; n.p_jm = p[id - nx]; n.p_km = p[id - sxy];
        if constexpr (WM == 2) {
            n.Vc = a.V[id];
            n.V_ip = a.V[id + 1]; n.V_im = a.V[id - 1]; n.V_jp = a.V[id + nx]; n.V_jm = a.V[id - nx];
            n.V_kp = a.V[id + sxy]; n.V_km = a.V[id - sxy];
            n.u_jp = u[id + nx]; n.v_jp = v[id + nx]; n.w_jp = w[id + nx];
            if constexpr (CEN) { n.u_jm = u[id - nx]; n.v_jm = v[id - nx]; n.w_jm = w[id - nx]; }
            if constexpr (TERRAIN) {
                n.u_kp = u[id + sxy]; n.v_kp = v[id + sxy]; n.w_kp = w[id + sxy];
                if constexpr (CEN) { n.u_km = u[id - sxy]; n.v_km = v[id - sxy]; n.w_km = w[id - sxy]; }
            }
        }
    }
}

// The layer terms of the pressure increment (air.cpp:2044-2292): the x, y and z slabs with
// their corner cross terms, added to r in the reference's order.  Reads the auxiliaries a.Qo.
// L: the slab axes compiled in (bit 0 x, bit 1 y, bit 2 z), for launches whose tiles can only
// meet those slabs; 7 compiles all.
template <typename T, bool TERRAIN, int L = 7>
__device__ __forceinline__ void dp_layers(const AirArgs<T> &a, int i, int j, int k, T mX, T mY, T mZ, T zX, T zY, T pc,
                                          T ddx, T ddy, T dwz, T &r)
{
    const Grid<T> &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t sxy = (size_t)nx * ny;
    const T dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, dt = g.dt, gP = a.gP;
    (void)sxy;
    const int n0 = a.qn[0], n1 = a.qn[1], n2 = a.qn[2], n3 = a.qn[3], n4 = a.qn[4], n5 = a.qn[5];
    // forward differences of a slab field along y or z with the metric, plus the cross-metric
    // z term over terrain: X-slab (stride n*ny per plane), Y-slab (nx*n), Z-slab (nx*ny)
    auto fY = [&](gptr<const T> Q, size_t q, size_t sy, size_t sz) {
        T v = (mY * (Q[q + sy] - Q[q])) * dyi;
        if constexpr (TERRAIN) v = v + (zY * (Q[q + sz] - Q[q])) * dzi;
        return v;
    };
    auto fX = [&](gptr<const T> Q, size_t q, size_t sz) {
        T v = (mX * (Q[q + 1] - Q[q])) * dxi;
        if constexpr (TERRAIN) v = v + (zX * (Q[q + sz] - Q[q])) * dzi;
        return v;
    };
    // X1 (air.cpp:2044-2057)
    if ((L & 1) && i < a.I1) {
        const T sx = a.x1.s[i], cx = a.x1.c[i];
        const size_t q = qxi(i, j, k, n0, ny), sz = (size_t)n0 * ny;
        gptr<const T> Qv = a.Qo[0][1], Qw = a.Qo[0][2];
        const T A = (cx * (gP * ddx)) * dt;
        const T B = (sx * (pc + gP * (fY(Qv, q, n0, sz) + ((mZ * (Qw[q + sz] - Qw[q])) * dzi)))) * dt;
        r += A - B;
    }
    // X2 (air.cpp:2059-2074)
    if ((L & 1) && i >= nx - a.I2) {
        const int ii = i - (nx - n1);
        const T sx = a.x2.s[i], cx = a.x2.c[i];
        const size_t q = qxi(ii, j, k, n1, ny), sz = (size_t)n1 * ny;
        gptr<const T> Qv = a.Qo[1][1], Qw = a.Qo[1][2];
        const T A = (cx * (gP * ddx)) * dt;
        const T B = (sx * (pc + gP * (fY(Qv, q, n1, sz) + ((mZ * (Qw[q + sz] - Qw[q])) * dzi)))) * dt;
        r += A - B;
    }
    // cross term of a y or z layer with an x layer (air.cpp:2091-2115 and the like):
    //   -(Qp_x*sx*so + gP*((1/bx-1)*so*(dQu_x/dX) + (1/bo-1)*sx*(dQ_o/dO)))*dt
    auto xcross_y = [&](T so, T dom) {
        if ((L & 1) && i < a.I1) {
            const T sx = a.x1.s[i], dxm = a.x1.d[i];
            const size_t q = qxi(i, j, k, n0, ny), sz = (size_t)n0 * ny;
            const T t1 = (dxm * so) * fX(a.Qo[0][0], q, sz);
            const T t2 = (dom * sx) * fY(a.Qo[0][1], q, n0, sz);
            r += (-((a.Qo[0][3][q] * sx) * so + gP * (t1 + t2))) * dt;
        } else if ((L & 1) && i >= nx - a.I2) {
            const int ii = i - (nx - n1);
            const T sx = a.x2.s[i], dxm = a.x2.d[i];
            const size_t q = qxi(ii, j, k, n1, ny), sz = (size_t)n1 * ny;
            const T t1 = (dxm * so) * fX(a.Qo[1][0], q, sz);
            const T t2 = (dom * sx) * fY(a.Qo[1][1], q, n1, sz);
            r += (-((a.Qo[1][3][q] * sx) * so + gP * (t1 + t2))) * dt;
        }
    };
    auto xcross_z = [&](T so, T dom) {
        if ((L & 1) && i < a.I1) {
            const T sx = a.x1.s[i], dxm = a.x1.d[i];
            const size_t q = qxi(i, j, k, n0, ny), sz = (size_t)n0 * ny;
            const T t1 = (dxm * so) * fX(a.Qo[0][0], q, sz);
            const T t2 = (dom * sx) * ((mZ * (a.Qo[0][2][q + sz] - a.Qo[0][2][q])) * dzi);
            r += (-((a.Qo[0][3][q] * sx) * so + gP * (t1 + t2))) * dt;
        } else if ((L & 1) && i >= nx - a.I2) {
            const int ii = i - (nx - n1);
            const T sx = a.x2.s[i], dxm = a.x2.d[i];
            const size_t q = qxi(ii, j, k, n1, ny), sz = (size_t)n1 * ny;
            const T t1 = (dxm * so) * fX(a.Qo[1][0], q, sz);
            const T t2 = (dom * sx) * ((mZ * (a.Qo[1][2][q + sz] - a.Qo[1][2][q])) * dzi);
            r += (-((a.Qo[1][3][q] * sx) * so + gP * (t1 + t2))) * dt;
        }
    };
    // Y1 (air.cpp:2076-2116)
    if ((L & 2) && j < a.J1) {
        const T sy = a.y1.s[j], cy = a.y1.c[j];
        const size_t q = qyi(i, j, k, nx, n2), sz = (size_t)nx * n2;
        gptr<const T> Qu = a.Qo[2][0], Qw = a.Qo[2][2];
        const T A = (cy * (gP * ddy)) * dt;
        const T B = (sy * (pc + gP * (fX(Qu, q, sz) + ((mZ * (Qw[q + sz] - Qw[q])) * dzi)))) * dt;
        r += A - B;
        xcross_y(sy, a.y1.d[j]);
    }
    // Y2 (air.cpp:2118-2158)
    if ((L & 2) && j >= ny - a.J2) {
        const int jj = j - (ny - n3);
        const T sy = a.y2.s[j], cy = a.y2.c[j];
        const size_t q = qyi(i, jj, k, nx, n3), sz = (size_t)nx * n3;
        gptr<const T> Qu = a.Qo[3][0], Qw = a.Qo[3][2];
        const T A = (cy * (gP * ddy)) * dt;
        const T B = (sy * (pc + gP * (fX(Qu, q, sz) + ((mZ * (Qw[q + sz] - Qw[q])) * dzi)))) * dt;
        r += A - B;
        xcross_y(sy, a.y2.d[j]);
    }
    // y-z cross term (air.cpp:2200-2224)
    auto ycross_z = [&](T sz_, T dzm) {
        if ((L & 2) && j < a.J1) {
            const T sy = a.y1.s[j], dym = a.y1.d[j];
            const size_t q = qyi(i, j, k, nx, n2), sz = (size_t)nx * n2;
            const T t1 = (dym * sz_) * fY(a.Qo[2][1], q, nx, sz);
            const T t2 = (dzm * sy) * ((mZ * (a.Qo[2][2][q + sz] - a.Qo[2][2][q])) * dzi);
            r += (-((a.Qo[2][3][q] * sy) * sz_ + gP * (t1 + t2))) * dt;
        } else if ((L & 2) && j >= ny - a.J2) {
            const int jj = j - (ny - n3);
            const T sy = a.y2.s[j], dym = a.y2.d[j];
            const size_t q = qyi(i, jj, k, nx, n3), sz = (size_t)nx * n3;
            const T t1 = (dym * sz_) * fY(a.Qo[3][1], q, nx, sz);
            const T t2 = (dzm * sy) * ((mZ * (a.Qo[3][2][q + sz] - a.Qo[3][2][q])) * dzi);
            r += (-((a.Qo[3][3][q] * sy) * sz_ + gP * (t1 + t2))) * dt;
        }
    };
    // z layer main terms: d/dX of Qu and d/dY of Qv in the z slab (air.cpp:2160-2292)
    auto zmain = [&](T sz_, T cz, gptr<const T> Qu, gptr<const T> Qv, size_t q) {
        const size_t sz = sxy;
        const T A = (((cz * a.gam) * a.Pav) * ((mZ * dwz) * dzi)) * dt;
        T inner;
        if constexpr (TERRAIN) {
            inner = (((((mX * (Qu[q + 1] - Qu[q])) * dxi) + ((zX * (Qu[q + sz] - Qu[q])) * dzi)) +
                      ((mY * (Qv[q + nx] - Qv[q])) * dyi)) + ((zY * (Qv[q + sz] - Qv[q])) * dzi));
        } else {
            inner = ((mX * (Qu[q + 1] - Qu[q])) * dxi) + ((mY * (Qv[q + nx] - Qv[q])) * dyi);
        }
        const T B = (sz_ * (pc + gP * inner)) * dt;
        r += A - B;
    };
    // Z1 (air.cpp:2160-2225)
    if ((L & 4) && k < a.K1) {
        const T sz = a.z1.s[k], cz = a.z1.c[k];
        zmain(sz, cz, a.Qo[4][0], a.Qo[4][1], qzi(i, j, k, nx, ny));
        xcross_z(sz, a.z1.d[k]);
        ycross_z(sz, a.z1.d[k]);
    }
    // Z2 (air.cpp:2227-2292)
    if ((L & 4) && k >= nz - a.K2) {
        const int kk = k - (nz - n5);
        const T sz = a.z2.s[k], cz = a.z2.c[k];
        zmain(sz, cz, a.Qo[5][0], a.Qo[5][1], qzi(i, j, kk, nx, ny));
        xcross_z(sz, a.z2.d[k]);
        ycross_z(sz, a.z2.d[k]);
    }
    (void)n4;
}

// The layer terms of the velocity increments (air.cpp:1579-1905): the x, y and z slabs with
// their corner cross terms, added to du, dv, dw in the reference's order.  Reads the pressure
// auxiliaries a.Qn[.][3] and the velocity auxiliaries a.Qo.
template <typename T, bool TERRAIN, int L = 7>
__device__ __forceinline__ void duvw_layers(const AirArgs<T> &a, int i, int j, int k, T mX, T mY, T mZ, T zX, T zY,
                                            T uc, T vc, T wc, T gpx, T gpy, T gpz, T &du, T &dv, T &dw)
{
    const Grid<T> &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t sxy = (size_t)nx * ny;
    const T dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, dt = g.dt, Aav = a.Aav;
    (void)sxy;
    const int n0 = a.qn[0], n1 = a.qn[1], n2 = a.qn[2], n3 = a.qn[3], n5 = a.qn[5];
    // backward differences of a slab field with the metric: d/dX (x and cross-z), d/dY
    // (y and cross-z); `sy`, `sz` are the slab strides in y and z
    auto bX = [&](gptr<const T> Q, size_t q, size_t sz) {
        T v = (mX * (Q[q] - Q[q - 1])) * dxi;
        if constexpr (TERRAIN) v = v + (zX * (Q[q] - Q[q - sz])) * dzi;
        return v;
    };
    auto bY = [&](gptr<const T> Q, size_t q, size_t sy, size_t sz) {
        T v = (mY * (Q[q] - Q[q - sy])) * dyi;
        if constexpr (TERRAIN) v = v + (zY * (Q[q] - Q[q - sz])) * dzi;
        return v;
    };
    // X1 / X2 (air.cpp:1579-1619)
    if ((L & 1) && i < a.I1) {
        const T sx = a.x1.s[i], cx = a.x1.c[i];
        const size_t q = qxi(i, j, k, n0, ny), sz = (size_t)n0 * ny;
        gptr<const T> Qp = a.Qn[0][3];
        du += ((-sx) * uc) * dt + (cx * (Aav * gpx)) * dt;
        dv += ((-sx) * (vc + Aav * bY(Qp, q, n0, sz))) * dt;
        dw += ((-sx) * (wc + Aav * ((mZ * (Qp[q] - Qp[q - sz])) * dzi))) * dt;
    }
    if ((L & 1) && i >= nx - a.I2) {
        const int ii = i - (nx - n1);
        const T sx = a.x2.s[i], cx = a.x2.c[i];
        const size_t q = qxi(ii, j, k, n1, ny), sz = (size_t)n1 * ny;
        gptr<const T> Qp = a.Qn[1][3];
        du += ((-sx) * uc) * dt + (cx * (Aav * gpx)) * dt;
        dv += ((-sx) * (vc + Aav * bY(Qp, q, n1, sz))) * dt;
        dw += ((-sx) * (wc + Aav * ((mZ * (Qp[q] - Qp[q - sz])) * dzi))) * dt;
    }
    // y layer with x cross terms (air.cpp:1621-1728)
    auto ylayer = [&](T sy, T cy, T dym, gptr<const T> Qp, size_t qy, int nq) {
        const size_t szy = (size_t)nx * nq;
        du += ((-sy) * (uc + Aav * bX(Qp, qy, szy))) * dt;
        dv += ((cy * Aav) * gpy) * dt - (sy * vc) * dt;
        dw += ((-sy) * (wc + Aav * ((mZ * (Qp[qy] - Qp[qy - szy])) * dzi))) * dt;
        if ((L & 1) && i < a.I1) {
            const T sx = a.x1.s[i], dxm = a.x1.d[i];
            const size_t q = qxi(i, j, k, n0, ny), sz = (size_t)n0 * ny;
            gptr<const T> Xp = a.Qn[0][3];
            du += (-((a.Qo[0][0][q] * sx) * sy + ((dxm * sy) * Aav) * bX(Xp, q, sz))) * dt;
            dv += (-((a.Qo[0][1][q] * sx) * sy + ((dym * sx) * Aav) * bY(Xp, q, n0, sz))) * dt;
            dw += (((-a.Qo[0][2][q]) * sx) * sy) * dt;
        } else if ((L & 1) && i >= nx - a.I2) {
            const int ii = i - (nx - n1);
            const T sx = a.x2.s[i], dxm = a.x2.d[i];
            const size_t q = qxi(ii, j, k, n1, ny), sz = (size_t)n1 * ny;
            gptr<const T> Xp = a.Qn[1][3];
            du += (-((a.Qo[1][0][q] * sx) * sy + ((dxm * sy) * Aav) * bX(Xp, q, sz))) * dt;
            dv += (-((a.Qo[1][1][q] * sx) * sy + ((dym * sx) * Aav) * bY(Xp, q, n1, sz))) * dt;
            dw += (((-a.Qo[1][2][q]) * sx) * sy) * dt;
        }
    };
    if ((L & 2) && j < a.J1) ylayer(a.y1.s[j], a.y1.c[j], a.y1.d[j], a.Qn[2][3], qyi(i, j, k, nx, n2), n2);
    if ((L & 2) && j >= ny - a.J2) ylayer(a.y2.s[j], a.y2.c[j], a.y2.d[j], a.Qn[3][3], qyi(i, j - (ny - n3), k, nx, n3), n3);
    // z layer with x and y cross terms (air.cpp:1730-1905)
    auto zlayer = [&](T sz, T cz, T dzm, gptr<const T> Qp, size_t qz) {
        du += ((-sz) * (uc + Aav * bX(Qp, qz, sxy))) * dt;
        dv += ((-sz) * (vc + Aav * bY(Qp, qz, nx, sxy))) * dt;
        dw += ((cz * Aav) * gpz) * dt - (sz * wc) * dt;
        if ((L & 1) && i < a.I1) {
            const T sx = a.x1.s[i], dxm = a.x1.d[i];
            const size_t q = qxi(i, j, k, n0, ny), s0 = (size_t)n0 * ny;
            gptr<const T> Xp = a.Qn[0][3];
            du += (-((a.Qo[0][0][q] * sx) * sz + ((dxm * sz) * Aav) * bX(Xp, q, s0))) * dt;
            dv += (((-a.Qo[0][1][q]) * sx) * sz) * dt;
            dw += (-((a.Qo[0][2][q] * sx) * sz + ((dzm * sx) * Aav) * ((mZ * (Xp[q] - Xp[q - s0])) * dzi))) * dt;
        } else if ((L & 1) && i >= nx - a.I2) {
            const int ii = i - (nx - n1);
            const T sx = a.x2.s[i], dxm = a.x2.d[i];
            const size_t q = qxi(ii, j, k, n1, ny), s1 = (size_t)n1 * ny;
            gptr<const T> Xp = a.Qn[1][3];
            du += (-((a.Qo[1][0][q] * sx) * sz + ((dxm * sz) * Aav) * bX(Xp, q, s1))) * dt;
            dv += (((-a.Qo[1][1][q]) * sx) * sz) * dt;
            dw += (-((a.Qo[1][2][q] * sx) * sz + ((dzm * sx) * Aav) * ((mZ * (Xp[q] - Xp[q - s1])) * dzi))) * dt;
        }
        if ((L & 2) && j < a.J1) {
            const T sy = a.y1.s[j], dym = a.y1.d[j];
            const size_t q = qyi(i, j, k, nx, n2), s2 = (size_t)nx * n2;
            gptr<const T> Yp = a.Qn[2][3];
            du += (((-a.Qo[2][0][q]) * sy) * sz) * dt;
            dv += (-((a.Qo[2][1][q] * sy) * sz + ((dym * sz) * Aav) * bY(Yp, q, nx, s2))) * dt;
            dw += (-((a.Qo[2][2][q] * sy) * sz + ((dzm * sy) * Aav) * ((mZ * (Yp[q] - Yp[q - s2])) * dzi))) * dt;
        } else if ((L & 2) && j >= ny - a.J2) {
            const int jj = j - (ny - n3);
            const T sy = a.y2.s[j], dym = a.y2.d[j];
            const size_t q = qyi(i, jj, k, nx, n3), s3 = (size_t)nx * n3;
            gptr<const T> Yp = a.Qn[3][3];
            du += (((-a.Qo[3][0][q]) * sy) * sz) * dt;
            dv += (-((a.Qo[3][1][q] * sy) * sz + ((dym * sz) * Aav) * bY(Yp, q, nx, s3))) * dt;
            dw += (-((a.Qo[3][2][q] * sy) * sz + ((dzm * sy) * Aav) * ((mZ * (Yp[q] - Yp[q - s3])) * dzi))) * dt;
        }
    };
    if ((L & 4) && k < a.K1) zlayer(a.z1.s[k], a.z1.c[k], a.z1.d[k], a.Qn[4][3], qzi(i, j, k, nx, ny));
    if ((L & 4) && k >= nz - a.K2) zlayer(a.z2.s[k], a.z2.c[k], a.z2.d[k], a.Qn[5][3], qzi(i, j, k - (nz - n5), nx, ny));
}

// ------------------------------------------------------------------------------------
// Pressure increment dt*R_p at one interior point: air::cal_fp (air.cpp:1953-2322).
// Reads u, v, w, p of the time level in g.un ... g.pn and the auxiliaries a.Qo.  Terms
// follow the reference's association; the TERRAIN terms are those of its nine-metric chain
// (air.cpp:1975-2292) that the separable grid multiplies by zero, inserted in chain order.
// ------------------------------------------------------------------------------------
// CEN (the RK4 path): the convective derivatives V d/dY ... are central differences instead
// of the reference's one-sided pair (backward for p, forward for u v w).  The one-sided pair
// advects the velocity components across the flow with a downwind difference, whose modes
// grow at up to 2|V|/dy; the central form is neutral.
// L: the slab axes the point may lie in (bit 0 x, 1 y, 2 z; 7 all, 0 none): the launches over
// the interior box and over the one-axis parts of the shell below know that the other
// slabs hold no point of theirs, so those layer terms, which would add nothing, are not
// compiled in.
template <typename T, int WM, bool TERRAIN, bool CEN = false, int L = 7>
__device__ __forceinline__ T air_dp(const AirArgs<T> &a, int i, int j, int k)
{
    const Grid<T> &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t sxy = (size_t)nx * ny;
    const size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    T mX, mY, mZ, zX = T(0), zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, mX, mY, mZ, zX, zY);
    else metric(g, i, j, k, mX, mY, mZ);
    const T dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, dt = g.dt, gP = a.gP;
    const T pc = g.pn[id];
    T ddx, ddy, dwz, r;
    if constexpr (WM != 1) {
        Nb<T> n;
        load_nb<T, WM, TERRAIN, CEN, true, false>(a, i, j, k, n);
        r = dp_core<T, WM, TERRAIN, CEN>(a, n, ddx, ddy, dwz);
    } else {
        const T dux = g.un[id + 1] - g.un[id];
        const T dvy = g.vn[id + nx] - g.vn[id];
        dwz = g.wn[id + sxy] - g.wn[id];
        T duz = T(0), dvz = T(0);
        T div;
        if constexpr (TERRAIN) {
            duz = g.un[id + sxy] - g.un[id];
            dvz = g.vn[id + sxy] - g.vn[id];
            div = (((((mX * dux) * dxi) + ((zX * duz) * dzi)) + ((mY * dvy) * dyi)) + ((zY * dvz) * dzi)) + ((mZ * dwz) * dzi);
        } else {
            div = (((mX * dux) * dxi) + ((mY * dvy) * dyi)) + ((mZ * dwz) * dzi);
        }
        // d/dX of u and d/dY of v (the damped terms of the x and y layers)
        ddx = (mX * dux) * dxi;
        ddy = (mY * dvy) * dyi;
        if constexpr (TERRAIN) {
            ddx = ddx + (zX * duz) * dzi;
            ddy = ddy + (zY * dvz) * dzi;
        }
            const T Uc = a.U[id], Vc = a.V[id], Wc = a.W[id];
            const T c1 = ((Uc * mX) * dxi) * dt, c2 = ((Vc * mY) * dyi) * dt;
            T c3;
            T UX = ((mX * (a.U[id + 1] - a.U[id - 1])) * dxi) * T(0.5);
            T VY = ((mY * (a.V[id + nx] - a.V[id - nx])) * dyi) * T(0.5);
            const T WZ = ((mZ * (a.W[id + sxy] - a.W[id - sxy])) * dzi) * T(0.5);
            if constexpr (TERRAIN) {
                c3 = ((((Uc * zX) + (Vc * zY)) + (Wc * mZ)) * dzi) * dt;
                UX = UX + ((zX * (a.U[id + sxy] - a.U[id - sxy])) * dzi) * T(0.5);
                VY = VY + ((zY * (a.V[id + sxy] - a.V[id - sxy])) * dzi) * T(0.5);
            } else {
                c3 = ((Wc * mZ) * dzi) * dt;
            }
            if constexpr (CEN)
                r = ((-c1) * ((g.pn[id + 1] - g.pn[id - 1]) * T(0.5)) - c2 * ((g.pn[id + nx] - g.pn[id - nx]) * T(0.5))) -
                    c3 * ((g.pn[id + sxy] - g.pn[id - sxy]) * T(0.5));
            else
                r = ((-c1) * (pc - g.pn[id - 1]) - c2 * (pc - g.pn[id - nx])) - c3 * (pc - g.pn[id - sxy]);
            r += (((-a.gam) * pc) * dt) * ((UX + VY) + WZ);
            r += (((-a.gam) * a.Pav) * dt) * div;
    }
    if constexpr (L != 0) dp_layers<T, TERRAIN, L>(a, i, j, k, mX, mY, mZ, zX, zY, pc, ddx, ddy, dwz, r);
    return r;
}

// Material id of the embedded porous obstacle at (i,j,k), 0 outside it.
template <typename T> __device__ __forceinline__ int obstacle_id(const AirArgs<T> &a, int i, int j, int k)
{
    if (!a.omask || i < a.ob[0] || i > a.ob[1] || j < a.ob[2] || j > a.ob[3] || k < a.ob[4] || k > a.ob[5]) return 0;
    const int ni = a.ob[1] - a.ob[0] + 1, nj = a.ob[3] - a.ob[2] + 1;
    return a.omask[(size_t)(i - a.ob[0]) + (size_t)ni * ((size_t)(j - a.ob[2]) + (size_t)nj * (k - a.ob[4]))];
}

// FB pressure update at one interior point: calculation_2D::Cal_pressure
// (calculation_2D.cpp:598-608) and Update_PML_Qp (air.cpp:2388-2439).  Writes p^{n+1} and
// the point's Q_p and returns p^{n+1}.
// L as for air_dp; below 7 the point is also outside the obstacle (the split launches run
// only without one).
// XF: the lanes on the x faces take the face node p = p(1 or nx - 2) and its Q_p, as
// rk_air_point (the Front / Back copy of calculation_2D.cpp:696-705 that the pressure face
// pass gives such a node).
template <typename T, bool WIND, bool TERRAIN = false, int L = 7, bool XF = false, bool OBST = true>
__device__ __forceinline__ T air_p_point(const AirArgs<T> &a, int i, int j, int k)
{
    const Grid<T> &g = a.g;
    const size_t id = (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
    if constexpr (L == 0) {
        const T pnew = g.pn[id] + air_dp<T, WIND ? 1 : 0, TERRAIN, false, 0>(a, i, j, k);
        g.pnn[id] = pnew;
        return pnew;
    }
    const bool face = XF && (i == 0 || i == g.nx - 1);
    const OwnQ<T> own = own_slabs<T, L>(a, i, j, k);
    T qp_old[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) qp_old[s] = own.in[s] ? a.Qo[s][3][own.idx[s]] : T(0);
    T r = T(0);
    if (!face) r = air_dp<T, WIND ? 1 : 0, TERRAIN, false, L>(a, i, j, k);
    if (const int mo = OBST && L == 7 && !XF ? obstacle_id(a, i, j, k) : 0) {
        // inside the obstacle the porous-medium continuity equation replaces the increment
        // (air.cpp:2294-2318)
        T mX, mY, mZ;
        metric(g, i, j, k, mX, mY, mZ);
        const size_t sxy = (size_t)g.nx * g.ny;
        const T div = (((mX * (g.un[id + 1] - g.un[id])) * g.dxi) + ((mY * (g.vn[id + g.nx] - g.vn[id])) * g.dyi)) +
                      ((mZ * (g.wn[id + sxy] - g.wn[id])) * g.dzi);
        r = (-a.ogama[mo]) * div;
    }
    T pnew = g.pn[id] + r;
    if constexpr (XF) {
        const int lane = (int)threadIdx.x;
        const T v = __shfl(pnew, i == 0 ? lane + 1 : i == g.nx - 1 ? lane - 1 : lane, 64);
        if (face) pnew = v;
    }
    g.pnn[id] = pnew;
    // Q_p += dt*p of this point; the value is final, face conditions only touch boundary nodes
#pragma unroll
    for (int s = 0; s < 6; ++s)
        if (own.in[s]) a.Qn[s][3][own.idx[s]] = qp_old[s] + g.dt * pnew;
    return pnew;
}

// ------------------------------------------------------------------------------------
// Velocity increments dt*R_u, dt*R_v, dt*R_w at one interior point: air::cal_fuvw
// (air.cpp:1476-1949).  Reads u, v, w of the time level in g.un ..., the pressure in g.pnn
// and the pressure auxiliaries in a.Qn[.][3] (FB: p^{n+1}; RK4: the stage state), the
// velocity auxiliaries in a.Qo.
// ------------------------------------------------------------------------------------
template <typename T, int WM, bool TERRAIN, bool CEN = false, int L = 7>
__device__ __forceinline__ void air_duvw(const AirArgs<T> &a, int i, int j, int k, T &du_, T &dv_, T &dw_)
{
    const Grid<T> &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t sxy = (size_t)nx * ny;
    const size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    T mX, mY, mZ, zX = T(0), zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, mX, mY, mZ, zX, zY);
    else metric(g, i, j, k, mX, mY, mZ);
    const T dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, dt = g.dt, Aav = a.Aav;
    gptr<const T> p = g.pnn;
    const T pc = p[id];
    const T uc = g.un[id], vc = g.vn[id], wc = g.wn[id];
    T gpx, gpy, gpz, du, dv, dw;
    if constexpr (WM != 1) {
        Nb<T> n;
        load_nb<T, WM, TERRAIN, CEN, false, true>(a, i, j, k, n);
        duvw_core<T, WM, TERRAIN, CEN>(a, n, gpx, gpy, gpz, du, dv, dw);
    } else {
        const T bpz = pc - p[id - sxy];
        gpx = (mX * (pc - p[id - 1])) * dxi;
        gpy = (mY * (pc - p[id - nx])) * dyi;
        gpz = (mZ * bpz) * dzi;
        if constexpr (TERRAIN) {
            gpx = gpx + (zX * bpz) * dzi;
            gpy = gpy + (zY * bpz) * dzi;
        }
            const T Uc = a.U[id], Vc = a.V[id], Wc = a.W[id];
            const T c1 = ((Uc * mX) * dxi) * dt, c2 = ((Vc * mY) * dyi) * dt;
            T c3;
            if constexpr (TERRAIN) c3 = ((((Uc * zX) + (Vc * zY)) + (Wc * mZ)) * dzi) * dt;
            else c3 = ((Wc * mZ) * dzi) * dt;
            gptr<const T> U = a.U, V = a.V, W = a.W;
            // central gradients of the mean flow: d/dX, d/dY, d/dZ
            auto gx = [&](gptr<const T> F) {
                T v = ((mX * (F[id + 1] - F[id - 1])) * dxi) * T(0.5);
                if constexpr (TERRAIN) v = v + ((zX * (F[id + sxy] - F[id - sxy])) * dzi) * T(0.5);
                return v;
            };
            auto gy = [&](gptr<const T> F) {
                T v = ((mY * (F[id + nx] - F[id - nx])) * dyi) * T(0.5);
                if constexpr (TERRAIN) v = v + ((zY * (F[id + sxy] - F[id - sxy])) * dzi) * T(0.5);
                return v;
            };
            auto gz = [&](gptr<const T> F) { return ((mZ * (F[id + sxy] - F[id - sxy])) * dzi) * T(0.5); };
            const T UX = gx(U), UY = gy(U), UZ = gz(U);
            const T VX = gx(V), VY = gy(V), VZ = gz(V);
            const T WX = gx(W), WY = gy(W), WZ = gz(W);
            // convective differences of one component: forward (reference) or central
            auto cv = [&](gptr<const T> F, T fc) {
                if constexpr (CEN)
                    return ((-c1) * ((F[id + 1] - F[id - 1]) * T(0.5)) - c2 * ((F[id + nx] - F[id - nx]) * T(0.5))) -
                           c3 * ((F[id + sxy] - F[id - sxy]) * T(0.5));
                else
                    return ((-c1) * (F[id + 1] - fc) - c2 * (F[id + nx] - fc)) - c3 * (F[id + sxy] - fc);
            };
            du = cv(g.un, uc);
            du += (((-uc) * UX) * dt - (vc * UY) * dt) - (wc * UZ) * dt;
            du += ((-Aav) * dt) * gpx;
            dv = cv(g.vn, vc);
            dv += (((-uc) * VX) * dt - (vc * VY) * dt) - (wc * VZ) * dt;
            dv += ((-Aav) * dt) * gpy;
            dw = cv(g.wn, wc);
            dw += (((-uc) * WX) * dt - (vc * WY) * dt) - (wc * WZ) * dt;
            dw += ((-Aav) * dt) * gpz;
    }
    if constexpr (L != 0) duvw_layers<T, TERRAIN, L>(a, i, j, k, mX, mY, mZ, zX, zY, uc, vc, wc, gpx, gpy, gpz, du, dv, dw);
    du_ = du; dv_ = dv; dw_ = dw;
}

// FB velocity update at one interior point with p at n+1: calculation_2D::Cal_velocity
// (calculation_2D.cpp:564-574) and Update_PML_Qvw (air.cpp:2323-2386).
// XF as for air_p_point: u = 0, v and w copied (calculation_2D.cpp:784-864).
template <typename T, bool WIND, bool TERRAIN = false, int L = 7, bool XF = false, bool OBST = true>
__device__ __forceinline__ void air_v_point(const AirArgs<T> &a, int i, int j, int k)
{
    const Grid<T> &g = a.g;
    const size_t id = (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
    if constexpr (L == 0) {
        T du, dv, dw;
        air_duvw<T, WIND ? 1 : 0, TERRAIN, false, 0>(a, i, j, k, du, dv, dw);
        g.unn[id] = g.un[id] + du;
        g.vnn[id] = g.vn[id] + dv;
        g.wnn[id] = g.wn[id] + dw;
        return;
    }
    const bool face = XF && (i == 0 || i == g.nx - 1);
    const OwnQ<T> own = own_slabs<T, L>(a, i, j, k);
    T qu_old[6], qv_old[6], qw_old[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        qu_old[s] = own.in[s] ? a.Qo[s][0][own.idx[s]] : T(0);
        qv_old[s] = own.in[s] ? a.Qo[s][1][own.idx[s]] : T(0);
        qw_old[s] = own.in[s] ? a.Qo[s][2][own.idx[s]] : T(0);
    }
    T du = T(0), dv = T(0), dw = T(0);
    if (!face) air_duvw<T, WIND ? 1 : 0, TERRAIN, false, L>(a, i, j, k, du, dv, dw);
    const T dt = g.dt;
    if (const int mo = OBST && L == 7 && !XF ? obstacle_id(a, i, j, k) : 0) {
        // inside the obstacle the porous-medium momentum equation replaces the increments
        // (air.cpp:1907-1945)
        T mX, mY, mZ;
        metric(g, i, j, k, mX, mY, mZ);
        gptr<const T> p = g.pnn;
        const T pc = p[id];
        const size_t sxy = (size_t)g.nx * g.ny;
        const T ib = a.oinvb[mo], f = ((T(-1.0) * ib) * dt) * a.oeff[mo];
        const T uc = g.un[id], vc = g.vn[id], wc = g.wn[id];
        du = (f * ((mX * (pc - p[id - 1])) * g.dxi) + uc * ib) - uc;
        dv = (f * ((mY * (pc - p[id - g.nx])) * g.dyi) + vc * ib) - vc;
        dw = (f * ((mZ * (pc - p[id - sxy])) * g.dzi) + wc * ib) - wc;
    }
    T un = g.un[id] + du, vn = g.vn[id] + dv, wn = g.wn[id] + dw;
    if constexpr (XF) {
        const int lane = (int)threadIdx.x, from = i == 0 ? lane + 1 : i == g.nx - 1 ? lane - 1 : lane;
        const T v1 = __shfl(vn, from, 64), w1 = __shfl(wn, from, 64);
        if (face) { un = T(0); vn = v1; wn = w1; }
    }
    g.unn[id] = un;
    g.vnn[id] = vn;
    g.wnn[id] = wn;
    // Q_u, Q_v, Q_w += dt*(u, v, w) of this point
#pragma unroll
    for (int s = 0; s < 6; ++s)
        if (own.in[s]) {
            a.Qn[s][0][own.idx[s]] = qu_old[s] + dt * un;
            a.Qn[s][1][own.idx[s]] = qv_old[s] + dt * vn;
            a.Qn[s][2][own.idx[s]] = qw_old[s] + dt * wn;
        }
}

// The two phases as separate passes, one thread per interior point.
template <typename T, bool WIND, bool TERRAIN>
__global__ __launch_bounds__(BX * BY) void k_air_p(const AirArgs<T> a)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i < 1 || i >= a.g.nx - 1 || j >= a.g.ny - 1) return;
    (void)air_p_point<T, WIND, TERRAIN>(a, i, j, k);
}

// Split forms of the two passes: a launch per box list, with the layer terms of the slab axes
// L its points can meet (0 for the interior box, one axis for the faces of the shell, 7 for
// its edges); together they cover the points of k_air_p / k_air_v with the same arithmetic.
// The split never runs with an embedded obstacle (split_regions refuses one), so the edge
// kernels are compiled without the obstacle branch.
// On the 512 x 256 x 128 grid with its 20-point layers (FP32) the interior launches take
// 39 + 54 us for 53 % of the points, the layer launches the rest of the 143 + 180 us.
template <typename T, bool WIND, bool TERRAIN, int L, int NB, bool XF = false>
__global__ __launch_bounds__(BX * BY) void k_air_p_box(const AirArgs<T> a, const BoxList<NB> l)
{
    int i, j, k;
    list_point(l, blockIdx.x, i, j, k);
    if (i < (XF ? 0 : 1) || i >= a.g.nx - (XF ? 0 : 1) || j >= a.g.ny - 1) return;
    (void)air_p_point<T, WIND, TERRAIN, L, XF, false>(a, i, j, k);
}
template <typename T, bool WIND, bool TERRAIN, int L, int NB, bool XF = false>
__global__ __launch_bounds__(BX * BY) void k_air_v_box(const AirArgs<T> a, const BoxList<NB> l)
{
    int i, j, k;
    list_point(l, blockIdx.x, i, j, k);
    if (i < (XF ? 0 : 1) || i >= a.g.nx - (XF ? 0 : 1) || j >= a.g.ny - 1) return;
    air_v_point<T, WIND, TERRAIN, L, XF, false>(a, i, j, k);
}

// The interior box of the split passes without a mean flow, in vector accesses: a lane takes
// P consecutive points of a row (P * sizeof(T) a multiple of 16 bytes), LT = BX * BY / P lanes
// one 64 x BY tile of the list, a block of 256 lanes 256 / LT tiles.  Every field the
// increments read at the points and at their j, k neighbours (j + 1, k + 1 for the pressure,
// j - 1, k - 1 for the velocity) is one vector load per lane; the i + 1 (i - 1) neighbour of
// the lane's last (first) point is a load of its own.  Each point's increment is dp_core /
// duvw_core over the same neighbour values as the one-point form (air_p_point /
// air_v_point with L = 0), so the bits are the same.  Rows must start aligned to the vector
// (fb_interior_vec: nx a multiple of P, the arrays' bases aligned).
template <typename T, int P> using VecP = T __attribute__((ext_vector_type(P)));
template <typename T, int P> __device__ __forceinline__ VecP<T, P> ldv(gptr<T> f, size_t q)
{
    return *(const __attribute__((address_space(1))) VecP<T, P> *)(f.p + q);
}
template <typename T, int P> __device__ __forceinline__ void stv(gptr<T> f, size_t q, const VecP<T, P> &v)
{
    *(__attribute__((address_space(1))) VecP<T, P> *)(f.p + q) = v;
}
// Point (i, j, k) of the lane's first point, false for a padding tile, a tile past the list's
// end or a row past the last interior one.
template <typename T, int P>
__device__ __forceinline__ bool vec_point(const AirArgs<T> &a, const BoxList<1> &l, int ntiles, int &i, int &j, int &k)
{
    constexpr int LR = BX / P;
    const int t = (int)(blockIdx.x * blockDim.y + threadIdx.y);
    if (t >= ntiles) return false;
    list_point(l, t, i, j, k);  // i = tile origin + threadIdx.x, j = first row + threadIdx.y + 1
    if (i < 0) return false;
    const int lane = (int)threadIdx.x, row = lane / LR;
    i += (lane - row * LR) * P - lane;
    j += row - (int)threadIdx.y;
    return j < a.g.ny - 1 && i >= 1 && i + P <= a.g.nx - 1;
}
// The pressure form runs on a separable grid only: over terrain (u and v at k + 1 and the
// per-column metric tables as well) the four-point pressure pass was slower than the
// one-point one, and the velocity pass faster.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_air_p_vec(const AirArgs<T> a, const BoxList<1> l, int ntiles)
{
    typedef VecP<T, P> V;
    int i, j, k;
    if (!vec_point<T, P>(a, l, ntiles, i, j, k)) return;
    const Grid<T> &g = a.g;
    const size_t nx = (size_t)g.nx, sxy = nx * g.ny, id = (size_t)i + nx * ((size_t)j + (size_t)g.ny * k);
    const V pc = ldv<T, P>(g.pn, id), uc = ldv<T, P>(g.un, id), vc = ldv<T, P>(g.vn, id), wc = ldv<T, P>(g.wn, id);
    const V vjp = ldv<T, P>(g.vn, id + nx), wkp = ldv<T, P>(g.wn, id + sxy);
    const T ulast = g.un[id + P];
    V out;
#pragma unroll
    for (int e = 0; e < P; ++e) {
        Nb<T> n;
        n.zX = T(0); n.zY = T(0);
        metric(g, i + e, j, k, n.mX, n.mY, n.mZ);
        n.pc = pc[e]; n.uc = uc[e]; n.vc = vc[e]; n.wc = wc[e];
        n.u_ip = e + 1 < P ? uc[(e + 1) & (P - 1)] : ulast;
        n.v_jp = vjp[e]; n.w_kp = wkp[e];
        T ddx, ddy, dwz;
        out[e] = pc[e] + dp_core<T, 0, false, false>(a, n, ddx, ddy, dwz);
    }
    stv<T, P>(g.pnn, id, out);
}
template <typename T, bool TERRAIN, int P>
__global__ __launch_bounds__(256) void k_air_v_vec(const AirArgs<T> a, const BoxList<1> l, int ntiles)
{
    typedef VecP<T, P> V;
    int i, j, k;
    if (!vec_point<T, P>(a, l, ntiles, i, j, k)) return;
    const Grid<T> &g = a.g;
    const size_t nx = (size_t)g.nx, sxy = nx * g.ny, id = (size_t)i + nx * ((size_t)j + (size_t)g.ny * k);
    const V pc = ldv<T, P>(g.pnn, id), pjm = ldv<T, P>(g.pnn, id - nx), pkm = ldv<T, P>(g.pnn, id - sxy);
    const V uc = ldv<T, P>(g.un, id), vc = ldv<T, P>(g.vn, id), wc = ldv<T, P>(g.wn, id);
    const T pfirst = g.pnn[id - 1];
    V uo, vo, wo;
#pragma unroll
    for (int e = 0; e < P; ++e) {
        Nb<T> n;
        n.zX = T(0); n.zY = T(0);
        if constexpr (TERRAIN) metric_terrain(g, i + e, j, k, n.mX, n.mY, n.mZ, n.zX, n.zY);
        else metric(g, i + e, j, k, n.mX, n.mY, n.mZ);
        n.uc = uc[e]; n.vc = vc[e]; n.wc = wc[e];
        n.pc = pc[e];
        n.p_im = e > 0 ? pc[(e + P - 1) & (P - 1)] : pfirst;
        n.p_jm = pjm[e]; n.p_km = pkm[e];
        T gpx, gpy, gpz, du, dv, dw;
        duvw_core<T, 0, TERRAIN, false>(a, n, gpx, gpy, gpz, du, dv, dw);
        uo[e] = uc[e] + du; vo[e] = vc[e] + dv; wo[e] = wc[e] + dw;
    }
    stv<T, P>(g.unn, id, uo); stv<T, P>(g.vnn, id, vo); stv<T, P>(g.wnn, id, wo);
}

template <typename T, bool WIND, bool TERRAIN>
__global__ __launch_bounds__(BX * BY) void k_air_v(const AirArgs<T> a)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i < 1 || i >= a.g.nx - 1 || j >= a.g.ny - 1) return;
    air_v_point<T, WIND, TERRAIN>(a, i, j, k);
}

// ------------------------------------------------------------------------------------
// Fused step: both phases in one pass.
//
// A block is FB_TY + 1 wavefronts.  Wavefront w holds row j = FB_TY*ty + w and lane l holds
// column i = 63*tx + l, so the block covers a 64 x (FB_TY+1) patch of which the 63 x FB_TY
// points with l >= 1 and w >= 1 are its own; lane 0 and wavefront 0 form the halo column
// and row.  The block walks a chunk of planes upwards.  On each plane every thread first
// computes p^{n+1} (and Q_p) at its point, then, after a barrier, the owned points compute
// their velocities, which read p^{n+1} and Q_p at (i-1, j, k), (i, j-1, k), (i, j, k-1):
// all written by this block before the barrier (halo points by the halo lanes, the plane
// below by the previous iteration or, at the start of a chunk, by an extra pressure-only
// pass over the plane below).  Halo values are recomputed from time level n and the old
// auxiliaries, which nothing modifies during the step, so they equal the owner's bit for
// bit.  Face nodes next to owned points (Front i = 0, West j = 0, South k = 0) get the
// value the face pass gives them -- a copy of the adjacent interior node or, on a porous
// South face, the porous plane Kp-2 through the interface; the face pass still runs
// afterwards over every boundary node and stores the same values.
// ------------------------------------------------------------------------------------
constexpr int FB_TY = 4;

// Q[f] of every active slab containing node (i,j,k) <- old + dt*val (the boundary-node part
// of air.cpp:2323-2439).
template <typename T> __device__ __forceinline__ void q_node(const AirArgs<T> &a, int f, int i, int j, int k, T val)
{
    const int nx = a.g.nx, ny = a.g.ny;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int n = a.qn[s];
        if (!n || (s == 2 && !a.y1on)) continue;
        const int ax = s >> 1, lo = a.qlo[s];
        const int c = ax == 0 ? i : ax == 1 ? j : k;
        if (c < lo || c >= lo + n) continue;
        const size_t q = ax == 0 ? qxi(i - lo, j, k, n, ny) : ax == 1 ? qyi(i, j - lo, k, nx, n) : qzi(i, j, k - lo, nx, ny);
        a.Qn[s][f][q] = a.Qo[s][f][q] + a.g.dt * val;
    }
}

// Workgroups are dealt to the 8 XCDs round-robin (b, b+8, b+16, ... share one XCD and its
// L2).  Give every XCD a contiguous run of tiles, so the halo a block re-reads was fetched
// into the same L2 by its neighbour.

template <typename T, bool WIND>
__global__ __launch_bounds__(64 * (FB_TY + 1)) void k_air_fb(const AirArgs<T> a, int ntx, int nty, int kc)
{
    const Grid<T> &g = a.g;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % ntx, ty = (tile / ntx) % nty, kq = tile / (ntx * nty);
    const int i = 63 * tx + (int)threadIdx.x, j = FB_TY * ty + (int)threadIdx.y;
    const int k0 = 1 + kq * kc, k1 = min(k0 + kc, nz - 1);
    const bool own = threadIdx.x >= 1 && threadIdx.y >= 1 && i <= nx - 2 && j <= ny - 2;
    // pressure task: the interior point (i, j), or the face node (0, j) / (i, 0) from its
    // interior neighbour; the edge (0, 0) is never read by the velocity update
    const bool face = (i == 0) != (j == 0);
    const bool ptask = i <= nx - 2 && j <= ny - 2 && !(i == 0 && j == 0);
    const int ip = i == 0 ? 1 : i, jp = j == 0 ? 1 : j;
    const size_t sxy = (size_t)nx * ny;
    // pressure (and Q_p) of plane k at this thread's point, with the face node it stands for
    auto pressure = [&](int k, bool faces) {
        // Keep the per-point address arithmetic here: hoisted out of the plane loop it
        // becomes one 64-bit induction pointer per array and the kernel runs out of registers.
        int ii = i, jj = j, ipp = ip, jpp = jp;
        asm volatile("" : "+v"(ii), "+v"(jj), "+v"(ipp), "+v"(jpp));
        if (!ptask || (face && !faces)) return;
        const T p = air_p_point<T, WIND, false, 7, false, false>(a, ipp, jpp, k);
        const size_t c = (size_t)ii + (size_t)nx * jj;
        if (face) {
            // Front (calculation_2D.cpp:696-705) or West (676-685) face node
            g.pnn[c + sxy * k] = p;
            q_node(a, 3, ii, jj, k, p);
        } else if (own && k == 1) {
            // South face k = 0: rigid copy, or the porous plane through the interface
            // (calculation_2D.cpp:686-695, 865-985)
            const T ps = a.south_pore ? a.pore_pnn[c + sxy * (a.pore_nz - 2)] : p;
            g.pnn[c] = ps;
            q_node(a, 3, ii, jj, 0, ps);
        }
    };
    // Software pipeline over the planes: one barrier per plane, after which the velocities
    // of plane k (reading planes k and k-1) and the pressures of plane k+1 are independent.
    if (k0 > 1) pressure(k0 - 1, false);  // the plane below the chunk: its own points only
    pressure(k0, true);
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
        __syncthreads();
        if (own) {
            int ii = i, jj = j;
            asm volatile("" : "+v"(ii), "+v"(jj));
            air_v_point<T, WIND, false, 7, false, false>(a, ii, jj, k);
        }
        if (k + 1 < k1) pressure(k + 1, true);
    }
}

// ------------------------------------------------------------------------------------
// RK4 (extension).  One kernel per stage and domain: the increment k = dt*R(stage state) at
// every interior point, then the register update
//   stage 0:   acc = x_n + b0 k,          dst = x_n + a0 k
//   stage 1,2: acc = acc + b_s k,         dst = x_n + a_s k
//   stage 3:                              dst = acc + b3 k   (= x_{n+1})
// (the FP64 form; FP32 below)
// with (a) = (1/2, 1/2, 1), (b) = (1/6, 1/3, 1/3, 1/6).  The PML auxiliaries integrate
// dQ/dt = f of the stage state with the same registers.  With COMP (FP32) acc sums only the
// increments b_s k_s and the last stage adds them to x_n with an exact TwoSum whose rounding
// error `err` is carried into the next step, so the state does not lose FP32 ulps every step
// (76 words per point and step instead of the 64 + 32 of compensating every stage sum).
// ------------------------------------------------------------------------------------

template <typename T> __device__ __forceinline__ T rk_a(int st) { return st == 0 ? T(0.5) : st == 1 ? T(0.5) : T(1.0); }
template <typename T> __device__ __forceinline__ T rk_b(int st)
{
    return (st == 0 || st == 3) ? T(1.0 / 6.0) : T(1.0 / 3.0);
}

// The register update on values: the caller reads the registers, so that a point's loads
// can all be issued before its first store (the field pointers may alias as far as the
// compiler knows, so a load after a store waits for it).
template <typename T, bool COMP>
__device__ __forceinline__ T rk_regs(int st, T xn, T k, T acc_in, T err_in, T &acc_out, T &err_out)
{
    if constexpr (!COMP) {
        (void)err_in; (void)err_out;
        if (st == 0) {
            acc_out = xn + rk_b<T>(0) * k;
            return xn + rk_a<T>(0) * k;
        }
        if (st < 3) {
            acc_out = acc_in + rk_b<T>(st) * k;
            return xn + rk_a<T>(st) * k;
        }
        return acc_in + rk_b<T>(3) * k;
    } else {
        if (st == 0) {
            acc_out = rk_b<T>(0) * k;
            return xn + rk_a<T>(0) * k;
        }
        if (st < 3) {
            acc_out = acc_in + rk_b<T>(st) * k;
            return xn + rk_a<T>(st) * k;
        }
        const T y = (acc_in + rk_b<T>(3) * k) + err_in;
        const T x = xn + y;
        const T bb = x - xn;
        err_out = (xn - (x - bb)) + (y - bb);  // TwoSum: xn + y = x + err exactly
        return x;
    }
}

// Pointer form of the register update, kept for the porous stage: there, reading each
// field's registers next to its update measured faster than reading them all first.
template <typename T, bool COMP>
__device__ __forceinline__ T rk_update(int st, T xn, T k, gptr<T> acc, gptr<T> lo, gptr<T> err, size_t id)
{
    if constexpr (!COMP) {
        if (st == 0) {
            acc[id] = xn + rk_b<T>(0) * k;
            return xn + rk_a<T>(0) * k;
        }
        if (st < 3) {
            acc[id] = acc[id] + rk_b<T>(st) * k;
            return xn + rk_a<T>(st) * k;
        }
        return acc[id] + rk_b<T>(3) * k;
    } else {
        // acc holds the step's increment sum b1 k1 + ... (small next to the state, so its FP32
        // rounding is small next to an ulp of the state); the last stage adds it and the
        // carried error of the state with an exact TwoSum and keeps the new rounding error
        (void)lo;
        if (st == 0) {
            acc[id] = rk_b<T>(0) * k;
            return xn + rk_a<T>(0) * k;
        }
        if (st < 3) {
            acc[id] = acc[id] + rk_b<T>(st) * k;
            return xn + rk_a<T>(st) * k;
        }
        const T y = (acc[id] + rk_b<T>(3) * k) + err[id];
        const T x = xn + y;
        const T bb = x - xn;
        err[id] = (xn - (x - bb)) + (y - bb);  // TwoSum: xn + y = x + err exactly
        return x;
    }
}

// Lawson (integrating-factor) RK4 for the porous velocities, x' = -(beta/dt) x + N(x) with
// k = dt N(stage): classical RK4 in the variable e^{beta t/dt} x, so the stiff Darcy drag is
// integrated exactly (no step limit from sigma dt / rho_e).  With cf = {E, E2, E2/2, E/6, E2/3},
// E = e^-beta, E2 = e^-beta/2:
//   stage 0: acc = E x_n + (E/6) k,   next = E2 x_n + (E2/2) k
//   stage 1: acc += (E2/3) k,         next = E2 x_n + k/2
//   stage 2: acc += (E2/3) k,         next = E x_n + E2 k
//   stage 3: x_{n+1} = acc + k/6
// COMP carries the low part and the state error as rk_regs does (the error scaled by E).
template <typename T, bool COMP>
__device__ __forceinline__ T rk_update_if(int st, T xn, T k, gptr<T> acc, gptr<T> lo, gptr<T> err, size_t id, gptr<const T> cf)
{
    const T E = cf[0], E2 = cf[1];
    const T w = st == 0 ? cf[3] : st < 3 ? cf[4] : T(1.0 / 6.0);
    const T cn = st == 2 ? E : E2, an = st == 0 ? cf[2] : st == 1 ? T(0.5) : E2;
    if constexpr (!COMP) {
        if (st == 0) {
            acc[id] = E * xn + w * k;
            return cn * xn + an * k;
        }
        if (st < 3) {
            acc[id] = acc[id] + w * k;
            return cn * xn + an * k;
        }
        return acc[id] + w * k;
    } else {
        const T y = w * k;
        const T a0 = st == 0 ? E * xn : acc[id];
        const T sm = a0 + y;
        const T bb = sm - a0;
        const T e = (a0 - (sm - bb)) + (y - bb);
        const T l = (st == 0 ? E * err[id] : lo[id]) + e;
        if (st < 3) {
            acc[id] = sm;
            lo[id] = l;
            return cn * xn + an * k;
        }
        const T x = sm + l;
        err[id] = l - (x - sm);
        return x;
    }
}

// Auxiliaries of the point's own slabs: dQ/dt = f (air.cpp:2323-2439 as an ODE) with the RK4
// register update, f the stage state xc at the point; each slab's reads before its stores.
template <typename T, int L>
__device__ __forceinline__ void rk_aux_point(const RKArgs<T> &r, int i, int j, int k, const T (&xc)[4])
{
    const AirArgs<T> &a = r.a;
    const int st = r.stage;
    const OwnQ<T> own = own_slabs<T, L>(a, i, j, k);
    const T dt = a.g.dt;
#pragma unroll
    for (int s = 0; s < 6; ++s)
        if (own.in[s]) {
            const size_t q = own.idx[s];
            T qx[4], qa[4] = {T(0), T(0), T(0), T(0)}, qao[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) qx[f] = r.qxn[s][f][q];
            if (st != 0) acc4_load(r.qacc[s][0], q, qa);
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                T unused;
                r.qdst[s][f][q] = rk_regs<T, false>(st, qx[f], dt * xc[f], qa[f], T(0), qao[f], unused);
            }
            if (st < 3) acc4_store(r.qacc[s][0], q, qao);
        }
}

// XF: the lanes on the x faces (i = 0 and nx - 1, rows and planes of the interior) take
// their face node of the destination state -- u = 0, v, w, p copied from the interior
// neighbour (calculation_2D.cpp:696-725, 784-864, the values the boundary stage's walks give
// such a node) -- and the node's auxiliaries, in place of the boundary stage: the lines of
// those nodes are streamed by this pass anyway, and one node per row costs a line per array
// in the boundary stage.  The lane is idle otherwise; the neighbour's values come across the
// wavefront (the face's neighbour lane is in the same 64-point row).
template <typename T, int WM, bool TERRAIN, bool COMP, int L = 7, bool XF = false>
__device__ __forceinline__ void rk_air_point(const RKArgs<T> &r, int i, int j, int k)
{
    const AirArgs<T> &a = r.a;
    const size_t id = (size_t)i + (size_t)a.g.nx * ((size_t)j + (size_t)a.g.ny * k);
    const int st = r.stage;
    const bool face = XF && (i == 0 || i == a.g.nx - 1);
    T kf[4] = {T(0), T(0), T(0), T(0)};
    if (!face) {
        kf[3] = air_dp<T, WM, TERRAIN, true, L>(a, i, j, k);
        air_duvw<T, WM, TERRAIN, true, L>(a, i, j, k, kf[0], kf[1], kf[2]);
    }
    const T xc[4] = {a.g.un[id], a.g.vn[id], a.g.wn[id], a.g.pn[id]};  // stage state at the point
    // every register read of the point first, then the arithmetic, then the stores
    T xn[4], ac[4] = {T(0), T(0), T(0), T(0)}, er[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        // level n: the stage state itself at stage 0; not needed by the last FP64 stage
        xn[f] = (st == 0 || face) ? xc[f] : (!COMP && st == 3) ? T(0) : r.xn[f][id];
        er[f] = (COMP && st == 3 && !face) ? r.err[f][id] : T(0);
    }
    if (st != 0 && !face) acc4_load(r.acc[0], id, ac);
    T d[4], ao[4], eo[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) d[f] = rk_regs<T, COMP>(st, xn[f], kf[f], ac[f], er[f], ao[f], eo[f]);
    if constexpr (XF) {
        const int lane = (int)threadIdx.x;
        const int from = i == 0 ? lane + 1 : i == a.g.nx - 1 ? lane - 1 : lane;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const T v = __shfl(d[f], from, 64);
            if (face) d[f] = f == 0 ? T(0) : v;
        }
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        r.dst[f][id] = d[f];
        if (COMP && st == 3 && !face) r.err[f][id] = eo[f];
    }
    if (st < 3 && !face) acc4_store(r.acc[0], id, ao);
    if constexpr (L != 0) rk_aux_point<T, L>(r, i, j, k, xc);
}

// The stage arguments (some 200 pointers) live in device memory: passed by value they
// would fill the scalar registers and spill to vector lanes at every use.
template <typename T, int WM, bool TERRAIN, bool COMP>
__global__ __launch_bounds__(BX * BY) void k_rk_air(const RKArgs<T> *__restrict__ rp)
{
    const RKArgs<T> &r = *rp;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i < 1 || i >= r.a.g.nx - 1 || j >= r.a.g.ny - 1) return;
    rk_air_point<T, WM, TERRAIN, COMP>(r, i, j, k);
}

// Tuned forms of the RK4 air stage: tiles of 64 x TY points and a flat grid.  COLUMNS orders
// the tiles plane-fastest (a tile's column of planes one after another) and hands each XCD a
// contiguous run of columns, so the planes k-1 and k+1 a tile reads were loaded into the same
// XCD's L2 by the tiles just before and after it; otherwise the order is plane by plane.
template <typename T, int WM, bool TERRAIN, bool COMP, int TY, bool COLUMNS>
__global__ __launch_bounds__(BX * TY) void k_rk_air_t(const RKArgs<T> *__restrict__ rp)
{
    const RKArgs<T> &r = *rp;
    const int nx = r.a.g.nx, ny = r.a.g.ny, K = r.a.g.nz - 2;
    const int tx = (nx + BX - 1) / BX, ty = (ny - 2 + TY - 1) / TY;
    int bx, by, k;
    if constexpr (COLUMNS) {
        const int t = xcd_tile(blockIdx.x, gridDim.x);
        k = t % K + 1;
        const int col = t / K;
        bx = col % tx;
        by = col / tx;
    } else {
        const int b = blockIdx.x;
        bx = b % tx;
        by = (b / tx) % ty;
        k = b / (tx * ty) + 1;
    }
    const int i = bx * BX + threadIdx.x, j = by * TY + threadIdx.y + 1;
    if (i < 1 || i >= nx - 1 || j >= ny - 1) return;
    rk_air_point<T, WM, TERRAIN, COMP>(r, i, j, k);
}

// Split form of the RK4 air stage (64 x 4 tiles), as k_air_p_box.  The FP32 interior kernel
// is held to 72 registers (7 waves per SIMD at least): left free it took 77 and ran 6; the
// FP64 one, and the FP32 one with a uniform mean flow (WM 1), would spill there.
template <typename T, int WM, bool TERRAIN, bool COMP, int L, int NB, bool XF = false>
__global__ __launch_bounds__(BX * BY) __attribute__((amdgpu_waves_per_eu((COMP && L == 0 && WM != 1) ? 7 : 1, 8))) void k_rk_air_box(
    const RKArgs<T> *__restrict__ rp, const BoxList<NB> l)
{
    const RKArgs<T> &r = *rp;
    int i, j, k;
    list_point(l, blockIdx.x, i, j, k);
    if (i < (XF ? 0 : 1) || i >= r.a.g.nx - (XF ? 0 : 1) || j >= r.a.g.ny - 1) return;
    rk_air_point<T, WM, TERRAIN, COMP, L, XF>(r, i, j, k);
}

// Interior RK4 air stage with the stage state staged through LDS.  Each block copies the
// fields its 64 x 4 tile reads -- u, v, w, p (and the mean flow V) on plane k with a halo of a
// 16-byte chunk in x and a row in y, and on planes k - 1, k + 1 the tile's own points -- into
// LDS with global_load_lds (16 bytes per lane, no registers held while the loads are in
// flight, so a block has its whole stencil input in flight at once), then every point takes
// its neighbours from LDS and runs the same dp_core / duvw_core and register update as
// rk_air_point: the same bits.  Mean flows of velocity_method 0 and 7; rows a multiple of 16
// bytes (nx % (16 / sizeof(T)) == 0), so every chunk is aligned.
template <typename T> struct Stage {
    static constexpr int H = 16 / (int)sizeof(T);   // elements per chunk, the x halo
    static constexpr int W = BX + 2 * H;            // a plane-k row: i0 - H .. i0 + 63 + H
    static constexpr int PK = (BY + 2) * W;         // plane k: rows j0 - 1 .. j0 + BY
    static constexpr int PS = BY * BX;              // planes k -+ 1: the tile's own points
    static constexpr int SZ = PK + 2 * PS;          // one field
    static constexpr int CK = PK / H, CS = PS / H;  // 16-byte chunks
    static constexpr int GK = (CK + 63) / 64, GS = CS / 64;  // wave loads per part
    static constexpr int G = GK + 2 * GS;           // wave loads per field
    static_assert(CS % 64 == 0, "side planes are whole wave loads");
};
typedef __attribute__((address_space(3))) void lds_void;

// Only grid points are loaded: chunks outside x = 0 .. nx - 1 or rows outside 0 .. ny - 1 (the
// planes k -+ 1: 0 .. ny - 2, the rows of points) are left out, which no point of the grid
// reads (tiles of the layer shell reach past the faces); the caller keeps nx a multiple of H.
template <typename T, int NF>
__device__ __forceinline__ void stage_fields(gptr<const T> const (&F)[5], T *lds, int nx, int ny, size_t sxy, int i0,
                                             int j0, int k)
{
    using S = Stage<T>;
    const int lane = threadIdx.x;
    // wave w (= threadIdx.y) issues the loads g = f * G + part with g % BY == w; the loops run
    // at compile time so that every field pointer is read at a constant index
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int part = 0; part < S::G; ++part) {
            if ((f * S::G + part) % BY != (int)threadIdx.y) continue;
            int c, limit, off, x, y, ylim;
            size_t e;
            if (part < S::GK) {  // plane k, row by row of W elements
                c = part * 64 + lane;
                const int r = c / (S::W / S::H), cc = c - r * (S::W / S::H);
                x = i0 - S::H + cc * S::H;
                y = j0 - 1 + r;
                ylim = ny;
                e = sxy * (size_t)k;
                limit = S::CK;
                off = f * S::SZ + part * 64 * S::H;
            } else {  // plane k - 1 (side 0) or k + 1 (side 1), the tile's rows
                const int q = part - S::GK, side = q / S::GS, gq = q - side * S::GS;
                c = gq * 64 + lane;
                const int r = c / (BX / S::H), cc = c - r * (BX / S::H);
                x = i0 + cc * S::H;
                y = j0 + r;
                ylim = ny - 1;
                e = sxy * (size_t)(side ? k + 1 : k - 1);
                limit = S::CS;
                off = f * S::SZ + S::PK + side * S::PS + gq * 64 * S::H;
            }
            if (x < 0 || x >= nx || y >= ylim) continue;
            e += (size_t)x + (size_t)nx * (size_t)y;
            if (c < limit)
                __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) const void *)(F[f] + e),
                                                 (lds_void *)(lds + off), 16, 0, 0);
        }
}

template <typename T, int WM, bool TERRAIN, bool COMP>
__global__ __launch_bounds__(BX * BY) void k_rk_air_lds(const RKArgs<T> *__restrict__ rp, const BoxList<1> l)
{
    static_assert(WM == 0 || WM == 2, "uniform mean flows take the register path");
    using S = Stage<T>;
    constexpr int NF = WM == 2 ? 5 : 4;
    __shared__ T lds[NF * S::SZ];
    const RKArgs<T> &r = *rp;
    const AirArgs<T> &a = r.a;
    const Grid<T> &g = a.g;
    int i, j, k;
    list_point(l, blockIdx.x, i, j, k);
    if (i < 1) return;  // a padding block (uniform over the block: no barrier is left waiting)
    const int ti = threadIdx.x, tj = threadIdx.y;
    const size_t nx = (size_t)g.nx, sxy = (size_t)g.nx * g.ny;
    const size_t id = (size_t)i + nx * ((size_t)j + (size_t)g.ny * k);
    {
        gptr<const T> const F[5] = {g.un, g.vn, g.wn, g.pn, a.V};
        stage_fields<T, NF>(F, lds, g.nx, g.ny, sxy, i - ti, j - tj, k);
    }
    // the point's registers while the stage state is in flight
    const int st = r.stage;
    T xn[4], ac[4] = {T(0), T(0), T(0), T(0)}, er[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        xn[f] = (st == 0 || (!COMP && st == 3)) ? T(0) : r.xn[f][id];
        er[f] = (COMP && st == 3) ? r.err[f][id] : T(0);
    }
    if (st != 0) acc4_load(r.acc[0], id, ac);
    Nb<T> n;
    n.zX = T(0); n.zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, n.mX, n.mY, n.mZ, n.zX, n.zY);
    else metric(g, i, j, k, n.mX, n.mY, n.mZ);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    auto K = [&](int f, int di, int dj) { return lds[f * S::SZ + (tj + 1 + dj) * S::W + (ti + S::H + di)]; };
    auto M = [&](int f) { return lds[f * S::SZ + S::PK + tj * BX + ti]; };
    auto P = [&](int f) { return lds[f * S::SZ + S::PK + S::PS + tj * BX + ti]; };
    n.uc = K(0, 0, 0); n.u_ip = K(0, 1, 0); n.u_jp = K(0, 0, 1); n.u_jm = K(0, 0, -1); n.u_kp = P(0); n.u_km = M(0);
    n.vc = K(1, 0, 0); n.v_jp = K(1, 0, 1); n.v_jm = K(1, 0, -1); n.v_kp = P(1); n.v_km = M(1);
    n.wc = K(2, 0, 0); n.w_jp = K(2, 0, 1); n.w_jm = K(2, 0, -1); n.w_kp = P(2); n.w_km = M(2);
    n.pc = K(3, 0, 0); n.p_im = K(3, -1, 0); n.p_jp = K(3, 0, 1); n.p_jm = K(3, 0, -1); n.p_kp = P(3); n.p_km = M(3);
    if constexpr (WM == 2) {
        n.Vc = K(4, 0, 0); n.V_ip = K(4, 1, 0); n.V_im = K(4, -1, 0); n.V_jp = K(4, 0, 1); n.V_jm = K(4, 0, -1);
        n.V_kp = P(4); n.V_km = M(4);
    }
    T kf[4], ddx, ddy, dwz, gpx, gpy, gpz;
    kf[3] = dp_core<T, WM, TERRAIN, true>(a, n, ddx, ddy, dwz);
    duvw_core<T, WM, TERRAIN, true>(a, n, gpx, gpy, gpz, kf[0], kf[1], kf[2]);
    const T xc[4] = {n.uc, n.vc, n.wc, n.pc};
    if (st == 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) xn[f] = xc[f];
    }
    T d[4], ao[4], eo[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) d[f] = rk_regs<T, COMP>(st, xn[f], kf[f], ac[f], er[f], ao[f], eo[f]);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        r.dst[f][id] = d[f];
        if (COMP && st == 3) r.err[f][id] = eo[f];
    }
    if (st < 3) acc4_store(r.acc[0], id, ao);
}

// The layer shell (or one part of it, L its slab axes) with the stage state staged through LDS
// as k_rk_air_lds stages the interior's: the same core on LDS operands, then the layer terms,
// the register update, the x-face lanes (XF) and the auxiliaries exactly as rk_air_point has
// them, so the bits are the same.  Every thread of a block stages and meets the barrier; the
// points outside the grid's interior rows leave after it.
template <typename T, int WM, bool TERRAIN, bool COMP, int L, int NB, bool XF>
__global__ __launch_bounds__(BX * BY) __attribute__((amdgpu_waves_per_eu(COMP ? 7 : 1, 8))) void k_rk_air_lds_box(const RKArgs<T> *__restrict__ rp, const BoxList<NB> l)
{
    static_assert(WM == 0 || WM == 2, "uniform mean flows take the register path");
    using S = Stage<T>;
    constexpr int NF = WM == 2 ? 5 : 4;
    __shared__ T lds[NF * S::SZ];
    const RKArgs<T> &r = *rp;
    const AirArgs<T> &a = r.a;
    const Grid<T> &g = a.g;
    int i, j, k;
    list_point(l, blockIdx.x, i, j, k);
    if (i < 0) return;  // a padding block (uniform over the block)
    const int ti = threadIdx.x, tj = threadIdx.y, nx = g.nx;
    const size_t sxy = (size_t)nx * g.ny;
    {
        gptr<const T> const F[5] = {g.un, g.vn, g.wn, g.pn, a.V};
        stage_fields<T, NF>(F, lds, nx, g.ny, sxy, i - ti, j - tj, k);
    }
    const bool live = i >= (XF ? 0 : 1) && i < nx - (XF ? 0 : 1) && j < g.ny - 1;
    const bool face = XF && (i == 0 || i == nx - 1);
    const size_t id = live ? (size_t)i + (size_t)nx * ((size_t)j + (size_t)g.ny * k) : 0;
    const int st = r.stage;
    T xn[4], ac[4] = {T(0), T(0), T(0), T(0)}, er[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        xn[f] = (!live || st == 0 || face || (!COMP && st == 3)) ? T(0) : r.xn[f][id];
        er[f] = (live && COMP && st == 3 && !face) ? r.err[f][id] : T(0);
    }
    if (live && st != 0 && !face) acc4_load(r.acc[0], id, ac);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!live) return;
    auto K = [&](int f, int di, int dj) { return lds[f * S::SZ + (tj + 1 + dj) * S::W + (ti + S::H + di)]; };
    auto M = [&](int f) { return lds[f * S::SZ + S::PK + tj * BX + ti]; };
    auto P = [&](int f) { return lds[f * S::SZ + S::PK + S::PS + tj * BX + ti]; };
    const T xc[4] = {K(0, 0, 0), K(1, 0, 0), K(2, 0, 0), K(3, 0, 0)};  // stage state at the point
    T kf[4] = {T(0), T(0), T(0), T(0)};
    if (!face) {
        Nb<T> n;
        n.zX = T(0); n.zY = T(0);
        if constexpr (TERRAIN) metric_terrain(g, i, j, k, n.mX, n.mY, n.mZ, n.zX, n.zY);
        else metric(g, i, j, k, n.mX, n.mY, n.mZ);
        n.uc = xc[0]; n.u_ip = K(0, 1, 0); n.u_jp = K(0, 0, 1); n.u_jm = K(0, 0, -1); n.u_kp = P(0); n.u_km = M(0);
        n.vc = xc[1]; n.v_jp = K(1, 0, 1); n.v_jm = K(1, 0, -1); n.v_kp = P(1); n.v_km = M(1);
        n.wc = xc[2]; n.w_jp = K(2, 0, 1); n.w_jm = K(2, 0, -1); n.w_kp = P(2); n.w_km = M(2);
        n.pc = xc[3]; n.p_im = K(3, -1, 0); n.p_jp = K(3, 0, 1); n.p_jm = K(3, 0, -1); n.p_kp = P(3); n.p_km = M(3);
        if constexpr (WM == 2) {
            n.Vc = K(4, 0, 0); n.V_ip = K(4, 1, 0); n.V_im = K(4, -1, 0); n.V_jp = K(4, 0, 1); n.V_jm = K(4, 0, -1);
            n.V_kp = P(4); n.V_km = M(4);
        }
        T ddx, ddy, dwz, gpx, gpy, gpz;
        kf[3] = dp_core<T, WM, TERRAIN, true>(a, n, ddx, ddy, dwz);
        if constexpr (L != 0) dp_layers<T, TERRAIN, L>(a, i, j, k, n.mX, n.mY, n.mZ, n.zX, n.zY, n.pc, ddx, ddy, dwz, kf[3]);
        duvw_core<T, WM, TERRAIN, true>(a, n, gpx, gpy, gpz, kf[0], kf[1], kf[2]);
        if constexpr (L != 0)
            duvw_layers<T, TERRAIN, L>(a, i, j, k, n.mX, n.mY, n.mZ, n.zX, n.zY, n.uc, n.vc, n.wc, gpx, gpy, gpz, kf[0],
                                       kf[1], kf[2]);
    }
    if (st == 0 || face) {
#pragma unroll
        for (int f = 0; f < 4; ++f) xn[f] = xc[f];
    }
    T d[4], ao[4], eo[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) d[f] = rk_regs<T, COMP>(st, xn[f], kf[f], ac[f], er[f], ao[f], eo[f]);
    if constexpr (XF) {
        const int from = i == 0 ? ti + 1 : i == nx - 1 ? ti - 1 : ti;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const T v = __shfl(d[f], from, 64);
            if (face) d[f] = f == 0 ? T(0) : v;
        }
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        r.dst[f][id] = d[f];
        if (COMP && st == 3 && !face) r.err[f][id] = eo[f];
    }
    if (st < 3 && !face) acc4_store(r.acc[0], id, ao);
    if constexpr (L != 0) rk_aux_point<T, L>(r, i, j, k, xc);
}

// Ensemble form: one launch steps every member.  The flat block index selects the member by
// the block prefix `first` (first[d] .. first[d+1]-1 are member d's BX x BY tiles, x fastest,
// then y, then interior planes) and reads that member's arguments from device memory.
__device__ __forceinline__ int batch_member(const int *first, int n, int b)
{
    int lo = 0, hi = n - 1;  // largest d with first[d] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Point of thread (threadIdx.x, threadIdx.y) in tile `local` of a member: tiles of 256
// consecutive columns c = i + nx*(j - 1) over the interior rows, plane after plane (a plane's
// tile count rounded up to a multiple of 8 when `pad`); false for a face column (i = 0,
// nx - 1) or past the last row (every point of a padding tile).
__device__ __forceinline__ bool batch_point(int nx, int ny, bool pad, int local, int &i, int &j, int &k)
{
    int per = (nx * (ny - 2) + kBatchBlock - 1) / kBatchBlock;
    if (pad) per = (per + 7) & ~7;
    const int c = (local % per) * kBatchBlock + threadIdx.x + BX * threadIdx.y;
    k = local / per + 1;
    j = c / nx + 1;
    i = c - (j - 1) * nx;
    return j < ny - 1 && i >= 1 && i < nx - 1;
}

// The batched stage also takes the x-face nodes of the interior rows whose interior neighbour
// is in the same wavefront (batch_takes_face), as the split single-domain launches do
// (rk_air_point with XF): the face lane runs the same code as the others and takes the
// neighbour's values across the wavefront.
template <typename T, int WM, bool TERRAIN, bool COMP>
__device__ __forceinline__ void rk_air_tile(const RKArgs<T> *__restrict__ args, const int *__restrict__ first, int n,
                                            int tile, bool pad)
{
    const int d = batch_member(first, n, tile);
    const RKArgs<T> &r = args[d];
    const int nx = r.a.g.nx, ny = r.a.g.ny;
    int i, j, k;
    const bool in = batch_point(nx, ny, pad, tile - first[d], i, j, k);
    if (in || (j < ny - 1 && batch_takes_face(nx, i, j))) rk_air_point<T, WM, TERRAIN, COMP, 7, true>(r, i, j, k);
}
template <typename T, int WM, bool TERRAIN, bool COMP, int FORM>
__global__ __launch_bounds__(BX * BY) void k_rk_air_batch(const RKArgs<T> *__restrict__ args, const int *__restrict__ first,
                                                         int n, int ntiles)
{
    if constexpr (FORM == 2) {
        // the tile index is uniform over the block, so every wave of it takes the same trips
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) rk_air_tile<T, WM, TERRAIN, COMP>(args, first, n, t, false);
    } else {
        rk_air_tile<T, WM, TERRAIN, COMP>(args, first, n, blockIdx.x, FORM == 1);
    }
}

// Porous layer: dt*R_p = -gama div u (porous.cpp:77-95); dt*R_u = -dt eff grad p, with the
// drag of porous.cpp:112-135, du/dt = ... - (sigma/rho_e) u, in the integrating factor.
template <typename T, bool TERRAIN, bool COMP>
__device__ __forceinline__ void rk_pore_point(const RKPoreArgs<T> &r, int i, int j, int k)
{
    const Grid<T> &g = r.g;
    const int nx = g.nx, ny = g.ny;
    const size_t sxy = (size_t)nx * ny;
    const size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    T mX, mY, mZ, zX = T(0), zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, mX, mY, mZ, zX, zY);
    else metric(g, i, j, k, mX, mY, mZ);
    const T dux = g.un[id + 1] - g.un[id], dvy = g.vn[id + nx] - g.vn[id], dwz = g.wn[id + sxy] - g.wn[id];
    T div;
    if constexpr (TERRAIN) {
        const T duz = g.un[id + sxy] - g.un[id], dvz = g.vn[id + sxy] - g.vn[id];
        div = (((((mX * g.dxi) * dux) + ((zX * g.dzi) * duz)) + ((mY * g.dyi) * dvy)) + ((zY * g.dzi) * dvz)) +
              ((mZ * g.dzi) * dwz);
    } else {
        div = (((mX * g.dxi) * dux) + ((mY * g.dyi) * dvy)) + ((mZ * g.dzi) * dwz);
    }
    const T pc = g.pn[id], bz = pc - g.pn[id - sxy];
    T gx = (mX * (pc - g.pn[id - 1])) * g.dxi, gy = (mY * (pc - g.pn[id - nx])) * g.dyi;
    const T gz = (mZ * bz) * g.dzi;
    if constexpr (TERRAIN) {
        gx = gx + (zX * bz) * g.dzi;
        gy = gy + (zY * bz) * g.dzi;
    }
    const T f = (-g.dt) * r.eff;
    const T xc[4] = {g.un[id], g.vn[id], g.wn[id], pc};
    const T kf[4] = {f * gx, f * gy, f * gz, (-r.gama) * div};
    const int st = r.stage;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const T xn = st == 0 ? xc[q] : r.xn[q][id];
        r.dst[q][id] = rk_update_if<T, COMP>(st, xn, kf[q], r.acc[q], r.lo[q], r.err[q], id, r.cf);
    }
    {
        const T xn = st == 0 ? xc[3] : (!COMP && st == 3) ? T(0) : r.xn[3][id];
        r.dst[3][id] = rk_update<T, COMP>(st, xn, kf[3], r.acc[3], r.lo[3], r.err[3], id);
    }
}

template <typename T, bool TERRAIN, bool COMP>
__global__ __launch_bounds__(BX * BY) void k_rk_pore(const RKPoreArgs<T> r)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
    if (i < 1 || i >= r.g.nx - 1 || j >= r.g.ny - 1) return;
    rk_pore_point<T, TERRAIN, COMP>(r, i, j, k);
}

template <typename T, bool TERRAIN, bool COMP>
__device__ __forceinline__ void rk_pore_tile(const RKPoreArgs<T> *__restrict__ args, const int *__restrict__ first, int n,
                                             int tile, bool pad)
{
    const int d = batch_member(first, n, tile);
    const RKPoreArgs<T> &r = args[d];
    const int nx = r.g.nx, ny = r.g.ny;
    int i, j, k;
    if (batch_point(nx, ny, pad, tile - first[d], i, j, k)) rk_pore_point<T, TERRAIN, COMP>(r, i, j, k);
}
template <typename T, bool TERRAIN, bool COMP, int FORM>
__global__ __launch_bounds__(BX * BY) void k_rk_pore_batch(const RKPoreArgs<T> *__restrict__ args,
                                                          const int *__restrict__ first, int n, int ntiles)
{
    if constexpr (FORM == 2) {
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) rk_pore_tile<T, TERRAIN, COMP>(args, first, n, t, false);
    } else {
        rk_pore_tile<T, TERRAIN, COMP>(args, first, n, blockIdx.x, FORM == 1);
    }
}

// ------------------------------------------------------------------------------------
// porous ground (porous.cpp:77-143) + FB add
// ------------------------------------------------------------------------------------
template <typename T, bool TERRAIN>
__global__ __launch_bounds__(BX * BY) void k_pore_p(const PoreArgs<T> a)
{
    const Grid<T> &g = a.g;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
    const int nx = g.nx, ny = g.ny;
    if (i < 1 || i >= nx - 1 || j >= ny - 1) return;
    const size_t sxy = (size_t)nx * ny;
    const size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    T mX, mY, mZ, zX = T(0), zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, mX, mY, mZ, zX, zY);
    else metric(g, i, j, k, mX, mY, mZ);
    T r;
    if constexpr (TERRAIN) {
        r = (-a.gama) * (((((mX * g.dxi) * (g.un[id + 1] - g.un[id])) + ((zX * g.dzi) * (g.un[id + sxy] - g.un[id]))) +
                          ((mY * g.dyi) * (g.vn[id + nx] - g.vn[id]))) + ((zY * g.dzi) * (g.vn[id + sxy] - g.vn[id])) +
                         ((mZ * g.dzi) * (g.wn[id + sxy] - g.wn[id])));
    } else {
        r = (-a.gama) * ((((mX * g.dxi) * (g.un[id + 1] - g.un[id])) + ((mY * g.dyi) * (g.vn[id + nx] - g.vn[id]))) +
                         ((mZ * g.dzi) * (g.wn[id + sxy] - g.wn[id])));
    }
    g.pnn[id] = g.pn[id] + r;
}

template <typename T, bool TERRAIN>
__global__ __launch_bounds__(BX * BY) void k_pore_v(const PoreArgs<T> a)
{
    const Grid<T> &g = a.g;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
    const int nx = g.nx, ny = g.ny;
    if (i < 1 || i >= nx - 1 || j >= ny - 1) return;
    const size_t sxy = (size_t)nx * ny;
    const size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    T mX, mY, mZ, zX = T(0), zY = T(0);
    if constexpr (TERRAIN) metric_terrain(g, i, j, k, mX, mY, mZ, zX, zY);
    else metric(g, i, j, k, mX, mY, mZ);
    gptr<const T> p = g.pnn;
    const T pc = p[id];
    const T f = (((T(-1.0) * a.invbeta) * g.dt) * a.eff);
    const T uc = g.un[id], vc = g.vn[id], wc = g.wn[id];
    const T bz = pc - p[id - sxy];
    T gx = (mX * (pc - p[id - 1])) * g.dxi, gy = (mY * (pc - p[id - nx])) * g.dyi;
    if constexpr (TERRAIN) {
        gx = gx + (zX * bz) * g.dzi;
        gy = gy + (zY * bz) * g.dzi;
    }
    const T du = (f * gx + uc * a.invbeta) - uc;
    const T dv = (f * gy + vc * a.invbeta) - vc;
    const T dw = (f * ((mZ * bz) * g.dzi) + wc * a.invbeta) - wc;
    g.unn[id] = uc + du;
    g.vnn[id] = vc + dv;
    g.wnn[id] = wc + dw;
}

// Frame shift along j (calculation_2D.cpp:461-477, air.cpp:2547-2573): out of place.
template <typename T>
__global__ void k_shift(T *dst, const T *src, int A, int B, int C, int tempJ, int lead, int s)
{
    const size_t total = (size_t)A * B * C;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)((t / A) % B);
        T v;
        if (j < tempJ + 1 && j + lead - 1 < B - 1) v = src[t + (size_t)A * s];
        else if (j >= tempJ + 1) v = T(0);
        else v = src[t];
        dst[t] = v;
    }
}

// RK4 frame shift of a y layer's auxiliaries [i + nx*(jj + n*k)], planes lo .. lo+n-1 of the
// frame: the field's shift rule of k_shift, reading the layer's own plane jj + s, or zero
// where that plane lies outside the layer.
template <typename T>
__global__ void k_shift_yslab(T *dst, const T *src, int nx, int n, int nz, int lo, int ny, int tempJ, int lead, int s)
{
    const size_t total = (size_t)nx * n * nz;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int jj = (int)((t / nx) % n), j = lo + jj;
        T v;
        if (j < tempJ + 1 && j + lead - 1 < ny - 1) v = jj + s < n ? src[t + (size_t)nx * s] : T(0);
        else if (j >= tempJ + 1) v = T(0);
        else v = src[t];
        dst[t] = v;
    }
}

template <typename T>
__global__ void k_copy_plane_j(T *p, int nx, int ny, int nz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= nx) return;
    const size_t b = (size_t)i + (size_t)nx * (size_t)ny * k;
    p[b] = p[b + nx];
}

// Ensemble receivers: receiver r reads member field p[r] at idx[r].
template <typename T>
__global__ void k_sample_batch(const T *const *p, const long long *idx, int n, double *out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const long long q = idx[r];
    double v = q >= 0 ? (double)p[r][q] : 0.0;
    if (fabs(v) < 1.0e-100) v = 0.0;
    out[r] = v;
}

template <typename T>
__global__ void k_sample(const T *p, const long long *idx, int n, double *out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const long long q = idx[r];
    double v = q >= 0 ? (double)p[q] : 0.0;
    if (fabs(v) < 1.0e-100) v = 0.0;
    out[r] = v;
}

// Contour subsample (calculation_2D::mpi_send_data, calculation_2D.cpp:1030-1105): the nodes
// with index % stride == 0 plus the last one on each axis, out[io + no_x*(jo + no_y*ko)].
template <typename T>
__global__ void k_subsample(const T *p, int nx, int ny, int nz, int stride, int ox, int oy, int oz, double *out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ox * oy * oz) return;
    const int io = (int)(t % ox), jo = (int)((t / ox) % oy), ko = (int)(t / ((long long)ox * oy));
    const int i = io < ox - 1 ? io * stride : nx - 1, j = jo < oy - 1 ? jo * stride : ny - 1, k = ko < oz - 1 ? ko * stride : nz - 1;
    out[t] = (double)p[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)];
}

template <typename T>
__global__ void k_finite(const T *f, size_t n, int *flag)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        if (!isfinite((double)f[t])) { atomicOr(flag, 1); return; }
}

// Time-series source, calculation_2D::UpdateSource (calculation_2D.cpp:371-385):
// p += amp * exp(-gc r^2) at every node, r from the node to the source
// (cal_InitialPressure, calculation_2D.cpp:292-298).  Coordinates in FP64.
template <typename T>
__global__ void k_add_source(T *p, int nx, int ny, int nz, const double *X, const double *Y, const double *Z,
                             double sx, double sy, double sz, double gc, double amp)
{
    const size_t n = (size_t)nx * ny * nz;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const size_t r = t / nx;
        const int j = (int)(r % ny), k = (int)(r / ny);
        const double x = X[i] - sx, y = Y[j] - sy, z = Z[k] - sz;
        const double rr = sqrt((x * x + y * y) + z * z);
        p[t] = p[t] + (T)(amp * (1.0 * exp((-gc) * rr * rr)));
    }
}

// Logarithmic wind of velocity_method 7 for one frame: V = u*/0.41 ln(zg/z0) along +y at the
// height zg above the local ground (V = 0 for zg <= z0), zg from the terrain height h of the
// frame (NULL: flat) and the stretching tk of the terrain-following grid.
template <typename T>
__global__ void k_wind_log(T *V, int nx, int ny, int nz, const double *h, const double *az, const double *tk,
                           double ustar, double z0)
{
    const size_t n = (size_t)nx * ny * nz;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t q = t % ((size_t)nx * ny);
        const int k = (int)(t / ((size_t)nx * ny));
        const double zk = h ? az[k] + h[q] * tk[k] : az[k];
        const double zb = h ? az[0] + h[q] * tk[0] : az[0];
        const double zg = zk - zb;
        V[t] = (T)(zg > z0 ? ustar / 0.41 * log(zg / z0) : 0.0);
    }
}

// Streaming copy used as the bandwidth reference and to calibrate HBM counters: one 16 B
// vector per thread, non-temporal (the data is touched once), as many workgroups as needed.
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy_words(v4f *dst, const v4f *src, size_t n4)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + t), dst + t);
}

// Many-stream reference: every lane reads one 4-byte word of each of NR arrays and writes one
// of each of NW arrays (sums of its reads), all reads in flight before the first store, as the
// stencil stages do with their fields: the bandwidth a kernel with that many streams can reach.
struct StreamPtrs { const float *in[kStreamMaxR]; float *out[kStreamMaxW]; };
template <int NR, int NW>
__global__ __launch_bounds__(256) void k_stream_words(StreamPtrs p, size_t n)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float a[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) a[r] = p.in[r][t];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        float s = 0.f;
#pragma unroll
        for (int r = w; r < NR; r += NW) s += a[r];
        p.out[w][t] = s;
    }
}

// Interior tiles: x tiles start at i = 0 so that a tile's rows cover whole 128-byte lines (the
// face threads i = 0 and i = nx - 1 return at once); y and z start at the first interior plane.
dim3 interior_grid(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny - 2 + BY - 1) / BY, nz - 2); }

// The split of a domain's tile space by the slab axes its tiles can meet.  Along each axis the
// tiles fall in three ranges: before, inside and after the range wholly outside that axis's
// layers (the terms of dp_layers / duvw_layers and the auxiliaries of own_slabs).  A box of
// three inside ranges is the interior (no layer terms), one with a single outside range meets
// only that axis's slabs, the rest (edges and corners of the shell) may meet any.  False when
// the interior is empty or an obstacle is embedded: the full launch then covers the domain.
struct Regions {
    BoxList<1> in;        // the interior
    BoxList<2> ax[3];     // the shell's parts that meet only x, only y, only z slabs
    BoxList<12> edge;     // the rest of the shell (edges and corners)
    BoxList<6> shell;     // the whole shell: planes below and above the interior, then rows
                          // before and after it, then columns left and right of it
};
template <int NB> void pack(BoxList<NB> &l, const std::vector<TileBox> &v, bool pad)
{
    l.n = (int)v.size();
    l.pad = pad ? 1 : 0;
    l.first[0] = 0;
    for (int q = 0; q < NB; ++q) {
        const int per = q < l.n ? v[q].nbx * v[q].nby : 0;
        l.b[q] = q < l.n ? v[q] : TileBox{0, 1, 0, 1, 1, 1};
        l.first[q + 1] = q < l.n ? l.first[q] + (pad ? (per + 7) & ~7 : per) * v[q].nk : INT_MAX;
    }
}
template <typename T> bool split_regions(const AirArgs<T> &a, Regions &R, bool pad)
{
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    if (a.omask) return false;
    int xa = std::max(1, a.I1), xb = std::min(nx - 1, nx - a.I2);
    int ya = std::max(1, a.J1), yb = std::min(ny - 1, ny - a.J2);
    int za = std::max(1, a.K1), zb = std::min(nz - 1, nz - a.K2);
    if (a.qn[0]) xa = std::max(xa, a.qlo[0] + a.qn[0]);
    if (a.qn[1]) xb = std::min(xb, a.qlo[1]);
    if (a.qn[2]) ya = std::max(ya, a.qlo[2] + a.qn[2]);
    if (a.qn[3]) yb = std::min(yb, a.qlo[3]);
    if (a.qn[4]) za = std::max(za, a.qlo[4] + a.qn[4]);
    if (a.qn[5]) zb = std::min(zb, a.qlo[5]);
    const int bx0 = (xa + BX - 1) / BX, bx1 = xb / BX;            // tiles [64 bx, 64 bx + 64) in [xa, xb)
    const int by0 = (ya - 1 + BY - 1) / BY, by1 = (yb - 1) / BY;  // rows [BY by + 1, BY by + BY + 1) in [ya, yb)
    if (bx1 <= bx0 || by1 <= by0 || zb <= za) return false;
    const int TX = (nx + BX - 1) / BX, TYt = (ny - 2 + BY - 1) / BY;
    // ranges (start, count) before / inside / after, per axis
    const int X[3][2] = {{0, bx0}, {bx0, bx1 - bx0}, {bx1, TX - bx1}};
    const int Y[3][2] = {{0, by0}, {by0, by1 - by0}, {by1, TYt - by1}};
    const int Z[3][2] = {{1, za - 1}, {za, zb - za}, {zb, nz - 1 - zb}};
    std::vector<TileBox> in, ax[3], edge, shell;
    auto add = [](std::vector<TileBox> &l, int x0, int nbx, int y0, int nby, int z0, int nk) {
        if (nbx > 0 && nby > 0 && nk > 0) l.push_back(TileBox{x0, nbx, y0, nby, z0, nk});
    };
    for (int cz = 0; cz < 3; ++cz)
        for (int cy = 0; cy < 3; ++cy) {
            const int outs = (cy != 1) + (cz != 1);
            if (outs == 2) {  // an edge for every x range: one box across x
                add(edge, 0, TX, Y[cy][0], Y[cy][1], Z[cz][0], Z[cz][1]);
                continue;
            }
            for (int cx = 0; cx < 3; ++cx) {
                const int o = outs + (cx != 1);
                std::vector<TileBox> &l = o == 0 ? in : o == 2 ? edge : cx != 1 ? ax[0] : cy != 1 ? ax[1] : ax[2];
                add(l, X[cx][0], X[cx][1], Y[cy][0], Y[cy][1], Z[cz][0], Z[cz][1]);
            }
        }
    const int nk = zb - za;
    add(shell, 0, TX, 0, TYt, 1, za - 1);
    add(shell, 0, TX, 0, TYt, zb, nz - 1 - zb);
    add(shell, 0, TX, 0, by0, za, nk);
    add(shell, 0, TX, by1, TYt - by1, za, nk);
    add(shell, 0, bx0, by0, by1 - by0, za, nk);
    add(shell, bx1, TX - bx1, by0, by1 - by0, za, nk);
    pack(R.in, in, pad);
    for (int q = 0; q < 3; ++q) pack(R.ax[q], ax[q], pad);
    pack(R.edge, edge, pad);
    pack(R.shell, shell, pad);
    return true;
}
template <int NB> unsigned list_blocks(const BoxList<NB> &l) { return (unsigned)l.first[l.n]; }


unsigned flat_blocks(size_t total)
{
    size_t b = (total + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b ? b : 1);
}

}  // namespace

// Whether the split FB passes take the x-face nodes of the interior rows (the interior is
// not empty and the East face lane has its neighbour in its own 64-point row); the face
// passes then walk only the x planes' edge rings.
template <typename T> bool fb_passes_take_xfaces(const AirArgs<T> &a)
{
    Regions R;
    return a.g.nx % BX != 1 && split_regions(a, R, false);
}

// Points per lane of the interior box's vector form (k_air_p_vec / k_air_v_vec), 0 for the
// one-point form.
template <typename T> constexpr int kFbVecP = sizeof(T) == 4 ? 4 : 0;
// Whether the interior box of the split passes runs the vector form: no mean flow, rows of a
// multiple of P points and every array of the two time levels aligned to the vector.
template <typename T> bool fb_interior_vec(const AirArgs<T> &a, bool wind)
{
    constexpr int P = kFbVecP<T>;
    if (P == 0) return false;
    const Grid<T> &g = a.g;
    const uintptr_t al = (uintptr_t)(T *)g.un | (uintptr_t)(T *)g.vn | (uintptr_t)(T *)g.wn | (uintptr_t)(T *)g.pn |
                         (uintptr_t)(T *)g.unn | (uintptr_t)(T *)g.vnn | (uintptr_t)(T *)g.wnn | (uintptr_t)(T *)g.pnn;
    return !wind && g.nx % P == 0 && al % (P * sizeof(T)) == 0;
}
template <typename T, bool TERRAIN, bool PRESSURE>
void launch_air_vec(const AirArgs<T> &a, const BoxList<1> &l, hipStream_t s)
{
    constexpr int P = kFbVecP<T> ? kFbVecP<T> : 1;
    constexpr int LT = BX * BY / P, TPB = 256 / LT;  // lanes per tile, tiles per block
    const int nt = (int)list_blocks(l);
    const dim3 grid((unsigned)((nt + TPB - 1) / TPB)), block(LT, TPB);
    if constexpr (PRESSURE) {
        static_assert(!TERRAIN, "the pressure pass's vector form runs on a separable grid");
        hipLaunchKernelGGL((k_air_p_vec<T, P>), grid, block, 0, s, a, l, nt);
    } else {
        hipLaunchKernelGGL((k_air_v_vec<T, TERRAIN, P>), grid, block, 0, s, a, l, nt);
    }
}

template <typename T> void launch_air_p(const AirArgs<T> &a, bool wind, hipStream_t s, bool pad, bool vec_ok)
{
    dim3 g = interior_grid(a.g.nx, a.g.ny, a.g.nz), b(BX, BY);
    const bool ter = a.g.tsh != nullptr;
    Regions R;
    if (split_regions(a, R, pad)) {
        // the pressure pass's vector form on a separable grid only
        const bool xf = fb_passes_take_xfaces<T>(a), vec = vec_ok && !ter && fb_interior_vec<T>(a, wind);
#define LEE_PS(W, TER)                                                                                       \
    do {                                                                                                     \
        if (R.in.n && vec) launch_air_vec<T, false, true>(a, R.in, s);                                          \
        else if (R.in.n) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 0, 1>), dim3(list_blocks(R.in)), b, 0, s, a, R.in); \
        if (R.ax[0].n && xf) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 1, 2, true>), dim3(list_blocks(R.ax[0])), b, 0, s, a, R.ax[0]); \
        else if (R.ax[0].n) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 1, 2>), dim3(list_blocks(R.ax[0])), b, 0, s, a, R.ax[0]); \
        if (R.ax[1].n) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 2, 2>), dim3(list_blocks(R.ax[1])), b, 0, s, a, R.ax[1]); \
        if (R.ax[2].n) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 4, 2>), dim3(list_blocks(R.ax[2])), b, 0, s, a, R.ax[2]); \
        if (R.edge.n && xf) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 7, 12, true>), dim3(list_blocks(R.edge)), b, 0, s, a, R.edge); \
        else if (R.edge.n) hipLaunchKernelGGL((k_air_p_box<T, W, TER, 7, 12>), dim3(list_blocks(R.edge)), b, 0, s, a, R.edge); \
    } while (0)
        if (ter) { if (wind) LEE_PS(true, true); else LEE_PS(false, true); }
        else { if (wind) LEE_PS(true, false); else LEE_PS(false, false); }
#undef LEE_PS
        return;
    }
    if (ter) {
        if (wind) hipLaunchKernelGGL((k_air_p<T, true, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_air_p<T, false, true>), g, b, 0, s, a);
    } else {
        if (wind) hipLaunchKernelGGL((k_air_p<T, true, false>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_air_p<T, false, false>), g, b, 0, s, a);
    }
}
template <typename T> void launch_air_v(const AirArgs<T> &a, bool wind, hipStream_t s, bool pad, bool vec_ok)
{
    dim3 g = interior_grid(a.g.nx, a.g.ny, a.g.nz), b(BX, BY);
    const bool ter = a.g.tsh != nullptr;
    Regions R;
    if (split_regions(a, R, pad)) {
        const bool xf = fb_passes_take_xfaces<T>(a), vec = vec_ok && fb_interior_vec<T>(a, wind);
#define LEE_VS(W, TER)                                                                                       \
    do {                                                                                                     \
        if (R.in.n && vec) launch_air_vec<T, TER, false>(a, R.in, s);                                           \
        else if (R.in.n) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 0, 1>), dim3(list_blocks(R.in)), b, 0, s, a, R.in); \
        if (R.ax[0].n && xf) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 1, 2, true>), dim3(list_blocks(R.ax[0])), b, 0, s, a, R.ax[0]); \
        else if (R.ax[0].n) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 1, 2>), dim3(list_blocks(R.ax[0])), b, 0, s, a, R.ax[0]); \
        if (R.ax[1].n) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 2, 2>), dim3(list_blocks(R.ax[1])), b, 0, s, a, R.ax[1]); \
        if (R.ax[2].n) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 4, 2>), dim3(list_blocks(R.ax[2])), b, 0, s, a, R.ax[2]); \
        if (R.edge.n && xf) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 7, 12, true>), dim3(list_blocks(R.edge)), b, 0, s, a, R.edge); \
        else if (R.edge.n) hipLaunchKernelGGL((k_air_v_box<T, W, TER, 7, 12>), dim3(list_blocks(R.edge)), b, 0, s, a, R.edge); \
    } while (0)
        if (ter) { if (wind) LEE_VS(true, true); else LEE_VS(false, true); }
        else { if (wind) LEE_VS(true, false); else LEE_VS(false, false); }
#undef LEE_VS
        return;
    }
    if (ter) {
        if (wind) hipLaunchKernelGGL((k_air_v<T, true, true>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_air_v<T, false, true>), g, b, 0, s, a);
    } else {
        if (wind) hipLaunchKernelGGL((k_air_v<T, true, false>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_air_v<T, false, false>), g, b, 0, s, a);
    }
}
template <typename T>
void launch_rk_air(const RKArgs<T> &r, const RKArgs<T> *rd, int wind, bool terrain, hipStream_t s)
{
    constexpr bool C = std::is_same<T, float>::value;
    dim3 g = interior_grid(r.a.g.nx, r.a.g.ny, r.a.g.nz), b(BX, BY);
#define LEE_RK(WM, TER) hipLaunchKernelGGL((k_rk_air<T, WM, TER, C>), g, b, 0, s, rd)
    if (terrain) {
        if (wind == 1) LEE_RK(1, true); else if (wind == 2) LEE_RK(2, true); else LEE_RK(0, true);
    } else {
        if (wind == 1) LEE_RK(1, false); else if (wind == 2) LEE_RK(2, false); else LEE_RK(0, false);
    }
#undef LEE_RK
}

// Variant v of the tuned stage launch (lee_ctx keeps the autotuner's choice): v = 0 the
// plane-by-plane 64x4 grid, 1..4 the flat grids of kRkTiles, 5 and 6 the split launches of
// split_regions: the interior and the shell, or the interior and the shell by slab axes; 7 and
// 8 as 5 and 6 with the interior staged through LDS, 9 and 10 with the shell staged as well
// (5 and 6 for a uniform mean flow or rows not a multiple of 16 bytes).  v = 0 when the
// interior is empty.
// Whether launch shape v of the stage takes the x-face nodes of the interior rows (the split
// shapes, when the interior is not empty and the East face lane has its neighbour in its own
// 64-point row); the boundary stage then walks only the x planes' edge rings.
template <typename T> bool rk_stage_takes_xfaces(const RKArgs<T> &r, int v)
{
    Regions R;
    return v >= 5 && v <= 10 && r.a.g.nx % BX != 1 && split_regions(r.a, R, true);
}

template <typename T>
void launch_rk_air_v(const RKArgs<T> &r, const RKArgs<T> *rd, int wind, bool terrain, int v, hipStream_t s)
{
    // the layer shell by slab axes: x, y, z layers and the edges (Regions R, block b, C, xf)
#define LEE_RK_AXES(WM, TER)                                                                                      \
    do {                                                                                                        \
        if (R.ax[0].n && xf) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 1, 2, true>), dim3(list_blocks(R.ax[0])), b, 0, s, rd, R.ax[0]); \
        else if (R.ax[0].n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 1, 2>), dim3(list_blocks(R.ax[0])), b, 0, s, rd, R.ax[0]); \
        if (R.ax[1].n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 2, 2>), dim3(list_blocks(R.ax[1])), b, 0, s, rd, R.ax[1]); \
        if (R.ax[2].n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 4, 2>), dim3(list_blocks(R.ax[2])), b, 0, s, rd, R.ax[2]); \
        if (R.edge.n && xf) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 7, 12, true>), dim3(list_blocks(R.edge)), b, 0, s, rd, R.edge); \
        else if (R.edge.n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 7, 12>), dim3(list_blocks(R.edge)), b, 0, s, rd, R.edge); \
    } while (0)
    Regions R;
    if (v <= 0 || (v >= 5 && !split_regions(r.a, R, true))) { launch_rk_air<T>(r, rd, wind, terrain, s); return; }
    const bool xf = rk_stage_takes_xfaces<T>(r, v);
    constexpr bool C = std::is_same<T, float>::value;
    if (v >= 7 && v <= 10 && wind != 1 && r.a.g.nx % Stage<T>::H == 0) {
        // the interior staged through LDS; the shell as v 5 (whole) or v 6 (by slab axes),
        // from registers (v 7, 8) or staged through LDS as well (v 9, 10)
        const dim3 b(BX, BY);
#define LEE_RKLS(WM, TER, LL, NB, LIST)                                                                            \
    do {                                                                                                        \
        if (LIST.n && xf) hipLaunchKernelGGL((k_rk_air_lds_box<T, WM, TER, C, LL, NB, true>), dim3(list_blocks(LIST)), b, 0, s, rd, LIST); \
        else if (LIST.n) hipLaunchKernelGGL((k_rk_air_lds_box<T, WM, TER, C, LL, NB, false>), dim3(list_blocks(LIST)), b, 0, s, rd, LIST); \
    } while (0)
#define LEE_RKL(WM, TER)                                                                                          \
    do {                                                                                                        \
        if (R.in.n) hipLaunchKernelGGL((k_rk_air_lds<T, WM, TER, C>), dim3(list_blocks(R.in)), b, 0, s, rd, R.in); \
        if (v == 9) { LEE_RKLS(WM, TER, 7, 6, R.shell); break; }                                                \
        if (v == 10) {                                                                                          \
            LEE_RKLS(WM, TER, 1, 2, R.ax[0]);                                                                   \
            if (R.ax[1].n) hipLaunchKernelGGL((k_rk_air_lds_box<T, WM, TER, C, 2, 2, false>), dim3(list_blocks(R.ax[1])), b, 0, s, rd, R.ax[1]); \
            if (R.ax[2].n) hipLaunchKernelGGL((k_rk_air_lds_box<T, WM, TER, C, 4, 2, false>), dim3(list_blocks(R.ax[2])), b, 0, s, rd, R.ax[2]); \
            LEE_RKLS(WM, TER, 7, 12, R.edge);                                                                   \
            break;                                                                                              \
        }                                                                                                       \
        if (v == 7) {                                                                                           \
            if (R.shell.n && xf) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 7, 6, true>), dim3(list_blocks(R.shell)), b, 0, s, rd, R.shell); \
            else if (R.shell.n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 7, 6>), dim3(list_blocks(R.shell)), b, 0, s, rd, R.shell); \
            break;                                                                                              \
        }                                                                                                       \
        LEE_RK_AXES(WM, TER);                                                                                   \
    } while (0)
        if (terrain) { if (wind == 2) LEE_RKL(2, true); else LEE_RKL(0, true); }
        else { if (wind == 2) LEE_RKL(2, false); else LEE_RKL(0, false); }
#undef LEE_RKL
#undef LEE_RKLS
        return;
    }
    if (v == 7 || v == 9) v = 5;
    if (v == 8 || v == 10) v = 6;
    if (v >= 5) {
        const dim3 b(BX, BY);
#define LEE_RKS(WM, TER)                                                                                          \
    do {                                                                                                        \
        if (R.in.n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 0, 1>), dim3(list_blocks(R.in)), b, 0, s, rd, R.in); \
        if (v == 5) {                                                                                           \
            if (R.shell.n && xf) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 7, 6, true>), dim3(list_blocks(R.shell)), b, 0, s, rd, R.shell); \
            else if (R.shell.n) hipLaunchKernelGGL((k_rk_air_box<T, WM, TER, C, 7, 6>), dim3(list_blocks(R.shell)), b, 0, s, rd, R.shell); \
            break;                                                                                              \
        }                                                                                                       \
        LEE_RK_AXES(WM, TER);                                                                                   \
    } while (0)
        if (terrain) {
            if (wind == 1) LEE_RKS(1, true); else if (wind == 2) LEE_RKS(2, true); else LEE_RKS(0, true);
        } else {
            if (wind == 1) LEE_RKS(1, false); else if (wind == 2) LEE_RKS(2, false); else LEE_RKS(0, false);
        }
#undef LEE_RKS
        return;
    }
    const int nx = r.a.g.nx, ny = r.a.g.ny, K = r.a.g.nz - 2;
    const int TYv = kRkTiles[v].ty;
    const unsigned nb = (unsigned)(((nx + BX - 1) / BX) * ((ny - 2 + TYv - 1) / TYv) * K);
#define LEE_RKT(TY, COL)                                                                              \
    do {                                                                                              \
        if (terrain) {                                                                                \
            if (wind == 1) hipLaunchKernelGGL((k_rk_air_t<T, 1, true, C, TY, COL>), dim3(nb), dim3(BX, TY), 0, s, rd);  \
            else if (wind == 2) hipLaunchKernelGGL((k_rk_air_t<T, 2, true, C, TY, COL>), dim3(nb), dim3(BX, TY), 0, s, rd); \
            else hipLaunchKernelGGL((k_rk_air_t<T, 0, true, C, TY, COL>), dim3(nb), dim3(BX, TY), 0, s, rd); \
        } else {                                                                                      \
            if (wind == 1) hipLaunchKernelGGL((k_rk_air_t<T, 1, false, C, TY, COL>), dim3(nb), dim3(BX, TY), 0, s, rd); \
            else if (wind == 2) hipLaunchKernelGGL((k_rk_air_t<T, 2, false, C, TY, COL>), dim3(nb), dim3(BX, TY), 0, s, rd); \
            else hipLaunchKernelGGL((k_rk_air_t<T, 0, false, C, TY, COL>), dim3(nb), dim3(BX, TY), 0, s, rd); \
        }                                                                                             \
    } while (0)
    switch (v) {
    case 1: LEE_RKT(4, true); break;
    case 2: LEE_RKT(8, true); break;
    case 3: LEE_RKT(8, false); break;
    default: LEE_RKT(16, true); break;
    }
#undef LEE_RKT
#undef LEE_RK_AXES
}

// Blocks of `kernel` (256 threads) the chip holds at once: the grid of a persistent form.
template <typename K> unsigned resident_blocks(K kernel)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, BX * BY, 0) != hipSuccess)
        return 0;
    return (unsigned)std::max(1, cus * per);
}
// One launch of a batched kernel in form `form` over `nblocks` tiles: a block per tile, or
// (form 2) the resident grid walking them.
template <typename K, typename... A>
void launch_form(K kernel, int form, int nblocks, hipStream_t s, A... args)
{
    unsigned g = (unsigned)nblocks;
    if (form == 2) {
        static const unsigned cap = resident_blocks(kernel);  // one per instantiation (one device)
        if (cap) g = std::min(g, cap);
    }
    hipLaunchKernelGGL(kernel, dim3(g), dim3(BX, BY), 0, s, args...);
}
template <typename T>
void launch_rk_air_batch(const RKArgs<T> *args, const int *first, int n, int nblocks, int wind, bool terrain, int form,
                         hipStream_t s)
{
    constexpr bool C = std::is_same<T, float>::value;
    if (nblocks <= 0) return;
#define LEE_RKF(WM, TER)                                                                                       \
    do {                                                                                                     \
        if (form == 2) launch_form(k_rk_air_batch<T, WM, TER, C, 2>, 2, nblocks, s, args, first, n, nblocks);    \
        else if (form == 1) launch_form(k_rk_air_batch<T, WM, TER, C, 1>, 1, nblocks, s, args, first, n, nblocks); \
        else launch_form(k_rk_air_batch<T, WM, TER, C, 0>, 0, nblocks, s, args, first, n, nblocks);             \
    } while (0)
    if (terrain) {
        if (wind == 1) LEE_RKF(1, true); else if (wind == 2) LEE_RKF(2, true); else LEE_RKF(0, true);
    } else {
        if (wind == 1) LEE_RKF(1, false); else if (wind == 2) LEE_RKF(2, false); else LEE_RKF(0, false);
    }
#undef LEE_RKF
}
template <typename T>
void launch_rk_pore_batch(const RKPoreArgs<T> *args, const int *first, int n, int nblocks, bool terrain, int form,
                          hipStream_t s)
{
    constexpr bool C = std::is_same<T, float>::value;
    if (nblocks <= 0) return;
#define LEE_RPF(TER)                                                                                           \
    do {                                                                                                     \
        if (form == 2) launch_form(k_rk_pore_batch<T, TER, C, 2>, 2, nblocks, s, args, first, n, nblocks);       \
        else if (form == 1) launch_form(k_rk_pore_batch<T, TER, C, 1>, 1, nblocks, s, args, first, n, nblocks);  \
        else launch_form(k_rk_pore_batch<T, TER, C, 0>, 0, nblocks, s, args, first, n, nblocks);                \
    } while (0)
    if (terrain) LEE_RPF(true); else LEE_RPF(false);
#undef LEE_RPF
}
template <typename T> void launch_sample_batch(const T *const *p, const long long *idx, int n, double *out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL((k_sample_batch<T>), dim3((n + 63) / 64), dim3(64), 0, s, p, idx, n, out);
}
template <typename T> void launch_rk_pore(const RKPoreArgs<T> &r, bool terrain, hipStream_t s)
{
    constexpr bool C = std::is_same<T, float>::value;
    dim3 g = interior_grid(r.g.nx, r.g.ny, r.g.nz), b(BX, BY);
    if (terrain) hipLaunchKernelGGL((k_rk_pore<T, true, C>), g, b, 0, s, r);
    else hipLaunchKernelGGL((k_rk_pore<T, false, C>), g, b, 0, s, r);
}

template <typename T> bool fused_fits(int nx) { return nx >= 3; }

template <typename T> void launch_air_fb(const AirArgs<T> &a, bool wind, hipStream_t s, int per_cu)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
    const int ntx = (nx - 2 + 62) / 63, nty = (ny - 2 + FB_TY - 1) / FB_TY;
    // split the planes into chunks until there are about per_cu blocks per CU
    int nch = std::max(1, std::min(nz - 2, (std::max(per_cu, 1) * cus + ntx * nty - 1) / (ntx * nty)));
    const int kc = (nz - 2 + nch - 1) / nch;
    nch = (nz - 2 + kc - 1) / kc;
    dim3 grid((unsigned)(ntx * nty * nch)), block(64, FB_TY + 1);
    if (wind) hipLaunchKernelGGL((k_air_fb<T, true>), grid, block, 0, s, a, ntx, nty, kc);
    else hipLaunchKernelGGL((k_air_fb<T, false>), grid, block, 0, s, a, ntx, nty, kc);
}

template <typename T> void launch_pore_p(const PoreArgs<T> &a, hipStream_t s)
{
    if (a.g.tsh) hipLaunchKernelGGL((k_pore_p<T, true>), interior_grid(a.g.nx, a.g.ny, a.g.nz), dim3(BX, BY), 0, s, a);
    else hipLaunchKernelGGL((k_pore_p<T, false>), interior_grid(a.g.nx, a.g.ny, a.g.nz), dim3(BX, BY), 0, s, a);
}
template <typename T> void launch_pore_v(const PoreArgs<T> &a, hipStream_t s)
{
    if (a.g.tsh) hipLaunchKernelGGL((k_pore_v<T, true>), interior_grid(a.g.nx, a.g.ny, a.g.nz), dim3(BX, BY), 0, s, a);
    else hipLaunchKernelGGL((k_pore_v<T, false>), interior_grid(a.g.nx, a.g.ny, a.g.nz), dim3(BX, BY), 0, s, a);
}
template <typename T> void launch_shift(T *dst, const T *src, int A, int B, int C, int tempJ, int lead, int sh, hipStream_t st)
{
    hipLaunchKernelGGL((k_shift<T>), dim3(flat_blocks((size_t)A * B * C)), dim3(256), 0, st, dst, src, A, B, C, tempJ, lead, sh);
}
template <typename T>
void launch_shift_yslab(T *dst, const T *src, int nx, int n, int nz, int lo, int ny, int tempJ, int lead, int sh,
                        hipStream_t st)
{
    hipLaunchKernelGGL((k_shift_yslab<T>), dim3(flat_blocks((size_t)nx * n * nz)), dim3(256), 0, st, dst, src, nx, n, nz, lo,
                       ny, tempJ, lead, sh);
}
template <typename T> void launch_copy_plane_j(T *p, int nx, int ny, int nz, hipStream_t s)
{
    hipLaunchKernelGGL((k_copy_plane_j<T>), dim3((nx + 63) / 64, nz), dim3(64), 0, s, p, nx, ny, nz);
}
template <typename T> void launch_sample(const T *p, const long long *idx, int n, double *out, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL((k_sample<T>), dim3((n + 63) / 64), dim3(64), 0, s, p, idx, n, out);
}
template <typename T> void launch_subsample(const T *p, int nx, int ny, int nz, int stride, int ox, int oy, int oz,
                                             double *out, hipStream_t s)
{
    const long long n = (long long)ox * oy * oz;
    if (n > 0) hipLaunchKernelGGL((k_subsample<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, nx, ny, nz, stride, ox, oy, oz, out);
}
template <typename T> void launch_finite_check(const T *f, size_t n, int *flag, hipStream_t s)
{
    hipLaunchKernelGGL((k_finite<T>), dim3(flat_blocks(n)), dim3(256), 0, s, f, n, flag);
}

template <typename T>
void launch_wind_log(T *V, int nx, int ny, int nz, const double *h, const double *az, const double *tk, double ustar,
                     double z0, hipStream_t s)
{
    hipLaunchKernelGGL((k_wind_log<T>), dim3(flat_blocks((size_t)nx * ny * nz)), dim3(256), 0, s, V, nx, ny, nz, h, az, tk,
                       ustar, z0);
}

template <typename T>
void launch_add_source(T *p, int nx, int ny, int nz, const double *X, const double *Y, const double *Z, double sx,
                       double sy, double sz, double gc, double amp, hipStream_t s)
{
    hipLaunchKernelGGL((k_add_source<T>), dim3(flat_blocks((size_t)nx * ny * nz)), dim3(256), 0, s, p, nx, ny, nz, X, Y,
                       Z, sx, sy, sz, gc, amp);
}

bool launch_stream_words(float *const *out, int nw, const float *const *in, int nr, size_t n, hipStream_t s)
{
    StreamPtrs p{};
    if (nr < 1 || nr > kStreamMaxR || nw < 1 || nw > kStreamMaxW) return false;
    for (int r = 0; r < nr; ++r) p.in[r] = in[r];
    for (int w = 0; w < nw; ++w) p.out[w] = out[w];
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
#define LEE_SW(R, W) if (nr == R && nw == W) { hipLaunchKernelGGL((k_stream_words<R, W>), g, b, 0, s, p, n); return true; }
    LEE_SW(1, 1) LEE_SW(4, 4) LEE_SW(8, 4) LEE_SW(12, 8) LEE_SW(20, 8) LEE_SW(28, 8)
#undef LEE_SW
    return false;
}

void launch_copy_words(float *dst, const float *src, size_t n, hipStream_t s)
{
    const size_t n4 = n / 4;
    hipLaunchKernelGGL(k_copy_words, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s, (v4f *)dst, (const v4f *)src, n4);
}

#define LEE_INST(T)                                                                                          \
    template void launch_air_p<T>(const AirArgs<T> &, bool, hipStream_t, bool, bool);                              \
    template void launch_air_v<T>(const AirArgs<T> &, bool, hipStream_t, bool, bool);                              \
    template bool fused_fits<T>(int);                                                                        \
    template void launch_rk_air<T>(const RKArgs<T> &, const RKArgs<T> *, int, bool, hipStream_t);            \
    template void launch_rk_pore<T>(const RKPoreArgs<T> &, bool, hipStream_t);                               \
    template void launch_rk_air_batch<T>(const RKArgs<T> *, const int *, int, int, int, bool, int, hipStream_t); \
    template void launch_rk_air_v<T>(const RKArgs<T> &, const RKArgs<T> *, int, bool, int, hipStream_t);     \
    template bool rk_stage_takes_xfaces<T>(const RKArgs<T> &, int);                                           \
    template bool fb_passes_take_xfaces<T>(const AirArgs<T> &);                                              \
    template void launch_rk_pore_batch<T>(const RKPoreArgs<T> *, const int *, int, int, bool, int, hipStream_t); \
    template void launch_sample_batch<T>(const T *const *, const long long *, int, double *, hipStream_t);     \
    template void launch_air_fb<T>(const AirArgs<T> &, bool, hipStream_t, int);                              \
    template void launch_pore_p<T>(const PoreArgs<T> &, hipStream_t);                                        \
    template void launch_pore_v<T>(const PoreArgs<T> &, hipStream_t);                                        \
    template void launch_shift<T>(T *, const T *, int, int, int, int, int, int, hipStream_t);                \
    template void launch_copy_plane_j<T>(T *, int, int, int, hipStream_t);                                   \
    template void launch_shift_yslab<T>(T *, const T *, int, int, int, int, int, int, int, int, hipStream_t);  \
    template void launch_sample<T>(const T *, const long long *, int, double *, hipStream_t);                \
    template void launch_finite_check<T>(const T *, size_t, int *, hipStream_t);                             \
    template void launch_wind_log<T>(T *, int, int, int, const double *, const double *, const double *,     \
                                     double, double, hipStream_t);                                           \
    template void launch_add_source<T>(T *, int, int, int, const double *, const double *, const double *,   \
                                       double, double, double, double, double, hipStream_t);                 \
    template void launch_subsample<T>(const T *, int, int, int, int, int, int, int, double *, hipStream_t);
LEE_INST(double)
LEE_INST(float)
#undef LEE_INST

}  // namespace lee
