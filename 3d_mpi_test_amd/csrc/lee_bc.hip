// lee_bc.hip -- all face conditions of one phase in one launch.
//
// The reference applies its faces one after another (pressure: N, E, W, F, B, then S or
// the porous chain S, E, W, F, B and the ghost-plane copy; velocity likewise with the
// East TDBC), each face copying from the plane next to it (calculation_2D.cpp:646-1028,
// airpore.cpp:324-398).  Where faces meet, a later face reads values an earlier face has
// just written.  Instead of one launch per face, each thread here takes one boundary node
// and evaluates its final value directly by walking the face order backwards: the value
// after faces 1..t is the last face (order <= t) containing the node applied to its
// source nodes as they were after faces 1..order-1.  The walk ends on interior nodes (the
// stencil output); it never needs a value the chain has not written yet, because every
// source of a face lies strictly inside along that face's axis and has no face that comes
// later in the order.  The arithmetic per node is the same as the sequential chain's, so
// results are bit-identical.
//
// Face order (1..6):  air       N E W F B  S (rigid) / I (ghost plane k=0 <- pore Kp-2)
//                     porous    S E W F B  I (ghost plane Kp-1 <- air k=1)
// The same launch also does the PML-auxiliary update of the boundary nodes, the East
// TDBC history update, and (pressure phase) the receiver samples of the step.
#include "lee_kernels.h"

namespace lee {

namespace {

template <typename T> __device__ __forceinline__ size_t gidx(const Grid<T> &g, int i, int j, int k)
{
    return (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
}

// Walking the chain backwards from face `t` to face 1, every face the node lies on either
// maps it to its source node (a copy from the plane next to the face: an index clamp),
// ends the walk with a value (a zero normal velocity, the TDBC velocity), or -- East
// pressure -- combines three source nodes, which are then only subject to face 1.
// The walk needs only the stencil output (interior nodes) from memory.

// ---- pressure -------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T east_extrap(const Grid<T> &g, int i, int k)
{
    const T a0 = g.pnn[gidx(g, i, g.ny - 2, k)], a1 = g.pnn[gidx(g, i, g.ny - 3, k)], a2 = g.pnn[gidx(g, i, g.ny - 4, k)];
    return T(15.0 / 8.0) * a0 - T(5.0 / 4.0) * a1 + T(3.0 / 8.0) * a2;
}

template <typename T> __device__ __forceinline__ T p_pore_walk(const BCArgs<T> &a, int i, int j, int k, int t);

// Ground impedance (extension): the plane k = 0 is the ground's interface plane, as the
// porous ground's top plane is when the air reads it through the interface copy
// (calculation_2D.cpp:878-886).  Its interior nodes are stencil output of the porous
// equations of the ground's top material (porous.cpp:77-143 with porous.cpp:44-50) -- the
// faces N E W F B then treat its edge nodes as on every other plane, and there is no South
// face.  The normal velocity (half a cell below the plane) follows the porous momentum
// equation with the pressure half a cell further down given by the impedance through the
// TDBC recursion of calculation_2D.cpp:986-1028 with K poles.  The plane's metric is the
// porous top plane's: gmx[i] = 1/X_over_x, gmy[j] = 1/Y_over_y, gmz = 1/Z_over_z of the plane
// k = 1, and on the terrain-following grid the cross terms zX, zY of the inverse Jacobian per
// node (the tables of gter) of a layer shifted rigidly with the ground, as porous.cpp's grid is
// under airpore.cpp's hills, not the air column's compressed metric.  The plane's z derivatives
// are one-sided towards k = 1.  With no hill every term reduces to the flat closure's bit for
// bit (the cross terms are zeros).
template <typename T> __device__ __forceinline__ T ground_p(const BCArgs<T> &a, int i, int j)
{
    const Grid<T> &g = a.air;
    const size_t e = gidx(g, i, j, 0);
    const T ex = (T)a.gmx[i], ey = (T)a.gmy[j], ez = (T)a.gmz;
    T div;
    if (a.gter) {
        const size_t c = (size_t)i + (size_t)g.nx * j, n2 = (size_t)g.nx * g.ny;
        const T zX = (T)a.gter[c], zY = (T)a.gter[n2 + c];
        const T dux = g.un[gidx(g, i + 1, j, 0)] - g.un[e], duz = g.un[gidx(g, i, j, 1)] - g.un[e];
        const T dvy = g.vn[gidx(g, i, j + 1, 0)] - g.vn[e], dvz = g.vn[gidx(g, i, j, 1)] - g.vn[e];
        div = ((ex * g.dxi * dux + zX * g.dzi * duz) + (ey * g.dyi * dvy + zY * g.dzi * dvz)) +
              ez * g.dzi * (g.wn[gidx(g, i, j, 1)] - g.wn[e]);
    } else {
        div = ex * g.dxi * (g.un[gidx(g, i + 1, j, 0)] - g.un[e]) + ey * g.dyi * (g.vn[gidx(g, i, j + 1, 0)] - g.vn[e]) +
              ez * g.dzi * (g.wn[gidx(g, i, j, 1)] - g.wn[e]);
    }
    return g.pn[e] + (-(T)a.g_gama * div);
}

// air pressure of a node as the stencil (or the ground plane) left it
template <typename T> __device__ __forceinline__ T p_air_base(const BCArgs<T> &a, int i, int j, int k)
{
    if (k == 0 && a.ground) return ground_p(a, i, j);
    return a.air.pnn[gidx(a.air, i, j, k)];
}

template <typename T> __device__ __forceinline__ T east_extrap_air(const BCArgs<T> &a, int i, int k)
{
    const Grid<T> &g = a.air;
    const T a0 = p_air_base(a, i, g.ny - 2, k), a1 = p_air_base(a, i, g.ny - 3, k), a2 = p_air_base(a, i, g.ny - 4, k);
    return T(15.0 / 8.0) * a0 - T(5.0 / 4.0) * a1 + T(3.0 / 8.0) * a2;
}

// air pressure after faces 1..t (t <= 6): N E W F B, then S or the ghost plane
template <typename T, bool ALLOW_PORE> __device__ __forceinline__ T p_air_walk(const BCArgs<T> &a, int i, int j, int k, int t)
{
    const Grid<T> &g = a.air;
    if (t >= 6 && k == 0 && !a.ground && (!a.porous || a.iface)) {
        if constexpr (ALLOW_PORE) {
            if (a.porous) return p_pore_walk(a, i, j, a.pore.nz - 2, 5);
        }
        k = 1;
    }
    if (t >= 5 && i == g.nx - 1) i = g.nx - 2;
    if (t >= 4 && i == 0) i = 1;
    if (t >= 3 && j == 0) j = 1;
    if (t >= 2 && j == g.ny - 1) return east_extrap_air(a, i, k == g.nz - 1 ? g.nz - 2 : k);
    if (t >= 1 && k == g.nz - 1) k = g.nz - 2;
    return p_air_base(a, i, j, k);
}

// porous pressure after faces 1..t: S E W F B, then the ghost plane Kp-1 <- air k=1
template <typename T> __device__ __forceinline__ T p_pore_walk(const BCArgs<T> &a, int i, int j, int k, int t)
{
    const Grid<T> &g = a.pore;
    if (t >= 6 && k == g.nz - 1 && a.iface) return p_air_walk<T, false>(a, i, j, 1, 5);
    if (t >= 5 && i == g.nx - 1) i = g.nx - 2;
    if (t >= 4 && i == 0) i = 1;
    if (t >= 3 && j == 0) j = 1;
    if (t >= 2 && j == g.ny - 1) return east_extrap(g, i, k == 0 ? 1 : k);
    if (t >= 1 && k == 0) k = 1;
    return g.pnn[gidx(g, i, j, k)];
}

// ---- velocity (component c: 0 u, 1 v, 2 w) --------------------------------------------
template <typename T> __device__ __forceinline__ gptr<T> vfield(const Grid<T> &g, int c) { return c == 0 ? g.unn : c == 1 ? g.vnn : g.wnn; }

// East TDBC velocity from the final boundary pressure and the history of the previous step
// (calculation_2D.cpp:1019-1028).
template <typename T> __device__ __forceinline__ double tdbc_v(const BCArgs<T> &a, int i, int k)
{
    const Grid<T> &g = a.air;
    const double *P = a.psi_old + ((size_t)i + (size_t)g.nx * k) * 6;
    double temp_psi = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) temp_psi += a.decay[q] * P[q];
    return a.inv * ((double)g.pnn[gidx(g, i, g.ny - 1, k)] / a.Zinf - temp_psi);
}

// Ground plane velocities of interior node (i, j) from the final pressures of the plane:
// tangential components by the porous momentum equation, the normal one (double: it feeds
// the FP64 histories) with the impedance's pressure half a cell below it,
//   w' = ib (w - cpl (p' - p_h)),  p_h = Zinf (1 + tb sum a) v + Zinf sum_k e^{-r_k tb} Psi_k,
// v = -w', solved for w'.
template <typename T> __device__ __forceinline__ double ground_w_of(const BCArgs<T> &a, int i, int j, double temp_psi)
{
    const Grid<T> &g = a.air;
    const size_t e = gidx(g, i, j, 0);
    const double hist = a.gZinf * temp_psi;
    return ((double)g.wn[e] - a.gcpl * ((double)g.pnn[e] - hist)) * a.g_ib / (1.0 + a.g_ib * a.gcpl * a.gzi);
}

// The ground histories are pole-major, Psi_q of node (i, j) at q * nx * ny + i + nx * j, so
// that each pole's reads and writes of a wave are contiguous.
template <typename T> __device__ __forceinline__ double ground_w(const BCArgs<T> &a, int i, int j)
{
    const size_t node = (size_t)i + (size_t)a.air.nx * j, nn = (size_t)a.air.nx * a.air.ny;
    double temp_psi = 0;
    for (int q = 0; q < a.gK; ++q) temp_psi += a.gdecay[q] * a.gpsi_old[q * nn + node];
    return ground_w_of(a, i, j, temp_psi);
}

template <typename T> __device__ __forceinline__ T ground_vel(const BCArgs<T> &a, int c, int i, int j)
{
    const Grid<T> &g = a.air;
    const size_t e = gidx(g, i, j, 0);
    const T ib = (T)a.g_ib, dt = a.dt, eff = (T)a.g_eff;
    // terrain: d/dX = ex d/dx + zX d/dz, d/dY = ey d/dy + zY d/dz with d/dz one-sided to
    // k = 1; on a flat grid the cross term is absent
    T pz = T(0), zX = T(0), zY = T(0);
    if (a.gter) {
        const size_t q = (size_t)i + (size_t)g.nx * j, n2 = (size_t)g.nx * g.ny;
        pz = (g.pnn[gidx(g, i, j, 1)] - g.pnn[e]) * g.dzi;
        zX = (T)a.gter[q];
        zY = (T)a.gter[n2 + q];
    }
    if (c == 0) {
        const T ex = (T)a.gmx[i];
        T gx = ex * (g.pnn[e] - g.pnn[gidx(g, i - 1, j, 0)]) * g.dxi;
        if (a.gter) gx = gx + zX * pz;
        return g.un[e] + (T(-1.0) * ib * dt * eff * gx + g.un[e] * ib - g.un[e]);
    }
    if (c == 1) {
        const T ey = (T)a.gmy[j];
        T gy = ey * (g.pnn[e] - g.pnn[gidx(g, i, j - 1, 0)]) * g.dyi;
        if (a.gter) gy = gy + zY * pz;
        return g.vn[e] + (T(-1.0) * ib * dt * eff * gy + g.vn[e] * ib - g.vn[e]);
    }
    return (T)ground_w(a, i, j);
}

template <typename T> __device__ __forceinline__ T v_pore_walk(const BCArgs<T> &a, int c, int i, int j, int k, int t);

// air velocity component c after faces 1..t
template <typename T, bool ALLOW_PORE> __device__ __forceinline__ T v_air_walk(const BCArgs<T> &a, int c, int i, int j, int k, int t)
{
    const Grid<T> &g = a.air;
    if (t >= 6 && k == 0 && !a.ground && (!a.porous || a.iface)) {
        if constexpr (ALLOW_PORE) {
            if (a.porous) return v_pore_walk(a, c, i, j, a.pore.nz - 2, 5);
        }
        if (c == 2) return T(0);
        k = 1;
    }
    if (t >= 5 && i == g.nx - 1) { if (c == 0) return T(0); i = g.nx - 2; }
    if (t >= 4 && i == 0) { if (c == 0) return T(0); i = 1; }
    if (t >= 3 && j == 0) { if (c == 1) return T(0); j = 1; }
    if (t >= 2 && j == g.ny - 1) {
        if (c == 1) return a.tdbc ? (T)tdbc_v(a, i, k) : T(0);
        j = g.ny - 2;
    }
    if (t >= 1 && k == g.nz - 1) { if (c == 2) return T(0); k = g.nz - 2; }
    if (k == 0 && a.ground) return ground_vel(a, c, i, j);
    return vfield(g, c)[gidx(g, i, j, k)];
}

template <typename T> __device__ __forceinline__ T v_pore_walk(const BCArgs<T> &a, int c, int i, int j, int k, int t)
{
    const Grid<T> &g = a.pore;
    if (t >= 6 && k == g.nz - 1 && a.iface) return v_air_walk<T, false>(a, c, i, j, 1, 5);
    if (t >= 5 && i == g.nx - 1) { if (c == 0) return T(0); i = g.nx - 2; }
    if (t >= 4 && i == 0) { if (c == 0) return T(0); i = 1; }
    if (t >= 3 && j == 0) { if (c == 1) return T(0); j = 1; }
    if (t >= 2 && j == g.ny - 1) { if (c == 1) return T(0); j = g.ny - 2; }
    if (t >= 1 && k == 0) { if (c == 2) return T(0); k = 1; }
    return vfield(g, c)[gidx(g, i, j, k)];
}

// i, j, k of boundary node t of an nx*ny*nz box: the two x planes, then the two y planes
// without their x edges, then the two z planes without x and y edges.  With xring the x
// planes contribute only their edge rings (rows j = 0 and ny - 1 of every plane k, then planes
// k = 0 and nz - 1 of the rows between): the nodes of the interior rows belong to the air
// stage (rk_stage_takes_xfaces).
__device__ __forceinline__ void boundary_node(long long t64, int nx, int ny, int nz, bool xring, int &i, int &j, int &k)
{
    // 32-bit divisions: the boundary counts fit (lee_create refuses grids whose counts do not).
    // Every branch sets all three coordinates as values: assigning a different output in each
    // branch let the optimiser merge the stores into one through a selected pointer, which put
    // i, j, k in private (scratch) memory.
    unsigned t = (unsigned)t64;
    const unsigned uy = (unsigned)ny, ux2 = (unsigned)(nx - 2), uz = (unsigned)nz;
    const unsigned nA = 2u * uy * uz, nB = 2u * ux2 * uz;
    int ii, jj, kk;
    bool ring = false;
    if (xring) {
        const unsigned per = 2u * uz + 2u * (uy - 2u);  // edge ring of one x plane
        if (t < 2u * per) {
            const unsigned side = t / per, r = t - side * per;
            const unsigned r2 = r < 2u * uz ? 0u : r - 2u * uz;
            ring = true;
            ii = side ? nx - 1 : 0;
            jj = r < 2u * uz ? ((r & 1u) ? ny - 1 : 0) : 1 + (int)(r2 >> 1);
            kk = r < 2u * uz ? (int)(r >> 1) : ((r2 & 1u) ? nz - 1 : 0);
        } else {
            t = t - 2u * per + nA;  // the y and z planes follow as without the ring
        }
    }
    if (ring) {
    } else if (t < nA) {
        const unsigned pl = uy * uz, side = t / pl, r = t - side * pl, q = r / uy;
        ii = side ? nx - 1 : 0; jj = (int)(r - q * uy); kk = (int)q;
    } else if (t < nA + nB) {
        t -= nA;
        const unsigned pl = ux2 * uz, side = t / pl, r = t - side * pl, q = r / ux2;
        ii = 1 + (int)(r - q * ux2); jj = side ? ny - 1 : 0; kk = (int)q;
    } else {
        t -= nA + nB;
        const unsigned pl = ux2 * (uy - 2u), side = t / pl, r = t - side * pl, q = r / ux2;
        ii = 1 + (int)(r - q * ux2); jj = 1 + (int)q; kk = side ? nz - 1 : 0;
    }
    i = ii; j = jj; k = kk;
}

// Q[f] <- Q[f] + dt*val[f] (new set from old) for the fields f0 .. f0+NF-1 of every PML slab
// containing air node (i,j,k) (air.cpp:2323-2439), in two phases so that the node's reads all
// precede its stores (a read after a store waits for it): qb_read before the node's field
// stores, qb_write after them.
template <typename T, int NF> struct QBnd {
    bool in[6];
    size_t q[6];
    T o[6][NF];
};
template <typename T, int NF>
__device__ __forceinline__ void qb_read(const BCArgs<T> &a, int f0, int i, int j, int k, QBnd<T, NF> &b)
{
    const Grid<T> &g = a.air;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int n = a.qn[s], ax = a.qaxis[s], lo = a.qlo[s];
        const int c = ax == 0 ? i : ax == 1 ? j : k;
        b.in[s] = n && !(s == 2 && !a.y1on) && c >= lo && c < lo + n;
        const int A = ax == 0 ? n : g.nx, B = ax == 1 ? n : g.ny;
        b.q[s] = (size_t)(ax == 0 ? i - lo : i) + (size_t)A * ((size_t)(ax == 1 ? j - lo : j) + (size_t)B * (ax == 2 ? k - lo : k));
        if (b.in[s]) {
#pragma unroll
            for (int f = 0; f < NF; ++f) b.o[s][f] = a.Qo[s][f0 + f][b.q[s]];
        }
    }
}
template <typename T, int NF>
__device__ __forceinline__ void qb_write(const BCArgs<T> &a, int f0, const QBnd<T, NF> &b, const T (&val)[NF])
{
#pragma unroll
    for (int s = 0; s < 6; ++s)
        if (b.in[s]) {
#pragma unroll
            for (int f = 0; f < NF; ++f) a.Qn[s][f0 + f][b.q[s]] = b.o[s][f] + a.dt * val[f];
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_p_all(const BCArgs<T> a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.nb_air) {
        int i, j, k;
        boundary_node(t, a.air.nx, a.air.ny, a.air.nz, a.xring == 1, i, j, k);
        const T v = p_air_walk<T, true>(a, i, j, k, 6);
        QBnd<T, 1> qb;
        qb_read<T, 1>(a, 3, i, j, k, qb);
        a.air.pnn[gidx(a.air, i, j, k)] = v;
        const T vals[1] = {v};
        qb_write<T, 1>(a, 3, qb, vals);
    } else if (t < a.nb_air + a.nb_pore) {
        int i, j, k;
        boundary_node(t - a.nb_air, a.pore.nx, a.pore.ny, a.pore.nz, false, i, j, k);
        a.pore.pnn[gidx(a.pore, i, j, k)] = p_pore_walk(a, i, j, k, 6);
    } else if (t < a.nb_air + a.nb_pore + a.nrec) {
        // receivers sample time level n, which this step does not modify (airpore.cpp:573-580)
        const int r = (int)(t - a.nb_air - a.nb_pore);
        const long long q = a.rec_idx[r];
        double v = q >= 0 ? (double)a.air.pn[q] : 0.0;
        if (fabs(v) < 1.0e-100) v = 0.0;
        a.rec_out[r] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bc_v_all(const BCArgs<T> a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.nb_air) {
        int i, j, k;
        boundary_node(t, a.air.nx, a.air.ny, a.air.nz, a.xring == 1, i, j, k);
        const size_t id = gidx(a.air, i, j, k);
        if (a.ground && k == 0 && i > 0 && i < a.air.nx - 1 && j > 0 && j < a.air.ny - 1) {
            // interior node of the ground plane: its velocities, then Update_Psi
            // (calculation_2D.cpp:1012-1017) with the velocity into the ground, reading the
            // histories once
            const T u = ground_vel(a, 0, i, j), v = ground_vel(a, 1, i, j);
            const size_t b = (size_t)i + (size_t)a.air.nx * j, nn = (size_t)a.air.nx * a.air.ny;
            double P[8];
            double temp_psi = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < a.gK) { P[q] = a.gpsi_old[q * nn + b]; temp_psi += a.gdecay[q] * P[q]; }
            const double wd = ground_w_of(a, i, j, temp_psi);
            const T w = (T)wd;
            QBnd<T, 3> qb;
            qb_read<T, 3>(a, 0, i, j, k, qb);
            a.air.unn[id] = u; a.air.vnn[id] = v; a.air.wnn[id] = w;
            const T vals[3] = {u, v, w};
            qb_write<T, 3>(a, 0, qb, vals);
            const double vg = -wd;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < a.gK) a.gpsi_new[q * nn + b] = a.gdecay[q] * P[q] + (a.ggain[q] * vg) * a.gtb;
            return;
        }
        const T u = v_air_walk<T, true>(a, 0, i, j, k, 6);
        const T v = v_air_walk<T, true>(a, 1, i, j, k, 6);
        const T w = v_air_walk<T, true>(a, 2, i, j, k, 6);
        QBnd<T, 3> qb;
        qb_read<T, 3>(a, 0, i, j, k, qb);
        a.air.unn[id] = u; a.air.vnn[id] = v; a.air.wnn[id] = w;
        const T vals[3] = {u, v, w};
        qb_write<T, 3>(a, 0, qb, vals);
        if (a.tdbc && j == a.air.ny - 1) {
            // history update with the TDBC velocity of this node (calculation_2D.cpp:1012-1017)
            const double vt = tdbc_v(a, i, k);
            const size_t b = ((size_t)i + (size_t)a.air.nx * k) * 6;
#pragma unroll
            for (int q = 0; q < 6; ++q) a.psi_new[b + q] = a.decay[q] * a.psi_old[b + q] + (a.gain[q] * vt) * a.tb;
        }
    } else if (t < a.nb_air + a.nb_pore) {
        int i, j, k;
        boundary_node(t - a.nb_air, a.pore.nx, a.pore.ny, a.pore.nz, false, i, j, k);
        const T u = v_pore_walk(a, 0, i, j, k, 6);
        const T v = v_pore_walk(a, 1, i, j, k, 6);
        const T w = v_pore_walk(a, 2, i, j, k, 6);
        const size_t id = gidx(a.pore, i, j, k);
        a.pore.unn[id] = u; a.pore.vnn[id] = v; a.pore.wnn[id] = w;
    }
}

// RK4 stage: auxiliaries of the boundary nodes (dQ/dt = f of the stage state, with the
// register update of k_rk_air) and the face values of the destination state.
template <typename T>
__device__ __forceinline__ T rk_q_regs(int st, T xn, T k, T acc_in, T &acc_out)
{
    const T b = (st == 0 || st == 3) ? T(1.0 / 6.0) : T(1.0 / 3.0);
    const T al = st == 0 ? T(0.5) : st == 1 ? T(0.5) : T(1.0);
    if (st == 0) { acc_out = xn + b * k; return xn + al * k; }
    if (st < 3) { acc_out = acc_in + b * k; return xn + al * k; }
    return acc_in + b * k;
}

template <typename T>
__device__ __forceinline__ void rk_bnd_node(const RKBndArgs<T> &r, long long t)
{
    const BCArgs<T> &a = r.b;
    if (t < a.nb_air) {
        int i, j, k;
        boundary_node(t, a.air.nx, a.air.ny, a.air.nz, a.xring == 1, i, j, k);
        if (a.xring == 2 && j > 0 && j < a.air.ny - 1 && k > 0 && k < a.air.nz - 1 && batch_takes_face(a.air.nx, i, j))
            return;  // the batched air stage took this face node
        const Grid<T> &g = a.air;
        const size_t id = gidx(g, i, j, k);
        // every read first -- the face walks, the node's stage state, the auxiliaries of each
        // PML slab holding the node -- then the stores: a read after a store waits for it (the
        // arrays may alias as far as the compiler knows), which made each slab a memory round
        // trip of its own
        const T p = p_air_walk<T, true>(a, i, j, k, 6);
        const T u = v_air_walk<T, true>(a, 0, i, j, k, 6);
        const T v = v_air_walk<T, true>(a, 1, i, j, k, 6);
        const T w = v_air_walk<T, true>(a, 2, i, j, k, 6);
        T sv[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) sv[f] = r.src[f][id];
        bool in[6];
        size_t qi[6];
        T qx[6][4], qa[6][4];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int n = a.qn[s], ax = a.qaxis[s], lo = a.qlo[s];
            const int c = ax == 0 ? i : ax == 1 ? j : k;
            in[s] = n && !(s == 2 && !a.y1on) && c >= lo && c < lo + n;
            const int A = ax == 0 ? n : g.nx, B = ax == 1 ? n : g.ny;
            qi[s] = (size_t)(ax == 0 ? i - lo : i) + (size_t)A * ((size_t)(ax == 1 ? j - lo : j) + (size_t)B * (ax == 2 ? k - lo : k));
            if (in[s]) {
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    qx[s][f] = r.qxn[s][f][qi[s]];
                    qa[s][f] = T(0);
                }
                if (r.stage != 0) acc4_load(r.qacc[s][0], qi[s], qa[s]);
            }
        }
        g.pnn[id] = p; g.unn[id] = u; g.vnn[id] = v; g.wnn[id] = w;
#pragma unroll
        for (int s = 0; s < 6; ++s)
            if (in[s]) {
                T ao[4];
#pragma unroll
                for (int f = 0; f < 4; ++f)
                    r.qdst[s][f][qi[s]] = rk_q_regs<T>(r.stage, qx[s][f], a.dt * sv[f], qa[s][f], ao[f]);
                if (r.stage < 3) acc4_store(r.qacc[s][0], qi[s], ao);
            }
    } else if (t < a.nb_air + a.nb_pore) {
        int i, j, k;
        boundary_node(t - a.nb_air, a.pore.nx, a.pore.ny, a.pore.nz, false, i, j, k);
        const Grid<T> &g = a.pore;
        const size_t id = gidx(g, i, j, k);
        const T p = p_pore_walk(a, i, j, k, 6);
        const T u = v_pore_walk(a, 0, i, j, k, 6);
        const T v = v_pore_walk(a, 1, i, j, k, 6);
        const T w = v_pore_walk(a, 2, i, j, k, 6);
        g.pnn[id] = p; g.unn[id] = u; g.vnn[id] = v; g.wnn[id] = w;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_rk_bnd(const RKBndArgs<T> *__restrict__ rp)
{
    rk_bnd_node<T>(*rp, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Ensemble form: blocks first[d] .. first[d+1]-1 walk member d's boundary nodes.
template <typename T>
__global__ __launch_bounds__(256) void k_rk_bnd_batch(const RKBndArgs<T> *__restrict__ args, const int *__restrict__ first,
                                                      int n)
{
    int lo = 0, hi = n - 1;
    const int b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid - 1;
    }
    rk_bnd_node<T>(args[lo], (long long)(b - first[lo]) * blockDim.x + threadIdx.x);
}

}  // namespace

template <typename T> void launch_rk_bnd(const RKBndArgs<T> &r, const RKBndArgs<T> *rd, hipStream_t s)
{
    const long long n = r.b.nb_air + r.b.nb_pore;
    hipLaunchKernelGGL((k_rk_bnd<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rd);
}
template <typename T> void launch_rk_bnd_batch(const RKBndArgs<T> *args, const int *first, int n, int nblocks, hipStream_t s)
{
    if (nblocks > 0) hipLaunchKernelGGL((k_rk_bnd_batch<T>), dim3((unsigned)nblocks), dim3(256), 0, s, args, first, n);
}
template void launch_rk_bnd_batch<float>(const RKBndArgs<float> *, const int *, int, int, hipStream_t);
template void launch_rk_bnd_batch<double>(const RKBndArgs<double> *, const int *, int, int, hipStream_t);
template void launch_rk_bnd<float>(const RKBndArgs<float> &, const RKBndArgs<float> *, hipStream_t);
template void launch_rk_bnd<double>(const RKBndArgs<double> &, const RKBndArgs<double> *, hipStream_t);

template <typename T> long long boundary_count(int nx, int ny, int nz, bool xring)
{
    if (nx <= 0) return 0;
    const long long xplanes = xring ? 2LL * (2LL * nz + 2LL * (ny - 2)) : 2LL * ny * nz;
    return xplanes + 2LL * (nx - 2) * nz + 2LL * (nx - 2) * (ny - 2);
}

template <typename T> void launch_bc_p_all(const BCArgs<T> &a, hipStream_t s)
{
    const long long n = a.nb_air + a.nb_pore + a.nrec;
    hipLaunchKernelGGL((k_bc_p_all<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
}

template <typename T> void launch_bc_v_all(const BCArgs<T> &a, hipStream_t s)
{
    const long long n = a.nb_air + a.nb_pore;
    hipLaunchKernelGGL((k_bc_v_all<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
}

template long long boundary_count<float>(int, int, int, bool);
template long long boundary_count<double>(int, int, int, bool);
template void launch_bc_p_all<float>(const BCArgs<float> &, hipStream_t);
template void launch_bc_p_all<double>(const BCArgs<double> &, hipStream_t);
template void launch_bc_v_all<float>(const BCArgs<float> &, hipStream_t);
template void launch_bc_v_all<double>(const BCArgs<double> &, hipStream_t);

}  // namespace lee
