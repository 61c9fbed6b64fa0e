// lee_kernels.h -- kernel argument blocks and launchers of the FB_p_v step on gfx950.
//
// Fields are SoA, x fastest: id = i + nx*(j + ny*k).  T is the state/arithmetic type
// (double for the FP64 path, float for the FP32 path).  PML auxiliary slabs use the
// x-fastest layout of their own box (x slab: ii + n*(j + ny*k); y slab: i + nx*(jj + n*k);
// z slab: i + nx*(j + ny*kk)).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

namespace lee {

// A device-memory pointer in the argument blocks.  The blocks live in device memory, so a
// pointer read from them is a generic (flat) address to the compiler: every access through it
// would be a flat load or store, which also counts against the scalar-load counter, so each
// wait for a pointer fetched from the block would wait for every field load in flight too.
// Indexing through gptr reads and writes through the global address space (global_load /
// global_store with a scalar base).  On the host it converts to and from T * like the raw
// pointer it holds.
template <typename T> struct gptr {
    T *p;
    gptr() = default;
    __host__ __device__ constexpr gptr(T *q) : p(q) {}
    __host__ __device__ constexpr gptr(std::nullptr_t) : p(nullptr) {}
    template <typename U, typename = typename std::enable_if<std::is_convertible<U *, T *>::value>::type>
    __host__ __device__ constexpr gptr(gptr<U> q) : p(q.p) {}
    __host__ __device__ constexpr operator T *() const { return p; }
    template <typename I> __device__ __forceinline__ __attribute__((address_space(1))) T &operator[](I i) const
    {
        return ((__attribute__((address_space(1))) T *)p)[i];
    }
};

// The RK4 accumulators of a PML slab's auxiliaries hold the four fields of an element side by
// side (element q at 4q .. 4q + 3: one 16-byte access per element in FP32, two in FP64); the
// stage arguments' qacc[s][f] all point at the slab's block.
template <typename T> __device__ __forceinline__ void acc4_load(gptr<T> base, size_t q, T (&a)[4])
{
    typedef T v4 __attribute__((ext_vector_type(4)));
    const v4 v = ((const __attribute__((address_space(1))) v4 *)base.p)[q];
    a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
}
template <typename T> __device__ __forceinline__ void acc4_store(gptr<T> base, size_t q, const T (&a)[4])
{
    typedef T v4 __attribute__((ext_vector_type(4)));
    ((__attribute__((address_space(1))) v4 *)base.p)[q] = v4{a[0], a[1], a[2], a[3]};
}

enum Face { F_N = 0, F_S = 1, F_E = 2, F_W = 3, F_F = 4, F_B = 5 };

// Per-index PML profile: sigma, (1 - 1/beta), (1/beta - 1).
template <typename T> struct DProf { gptr<const T> s, c, d; };

template <typename T> struct Grid {
    int nx, ny, nz;
    gptr<T> un, vn, wn, pn;  // time level n
    gptr<T> unn, vnn, wnn, pnn;  // time level n+1
    // FP64: X_over_x, Y_over_y, Z_over_z (the Jacobian is formed per point as the
    // reference does).  FP32: the reciprocals 1/X_over_x ... (= the separable metric).
    gptr<const T> mx, my, mz;
    T dxi, dyi, dzi, dt;
    // terrain-following grid (extension): Z_over_z = mz[k]*tsh[c], Z_over_x = thx[c]*ttk[k],
    // Z_over_y = thy[c]*ttk[k], c = i + nx*j; mx, my, mz then hold X_over_x ... in both precisions
    gptr<const T> tsh, thx, thy, ttk;
};

template <typename T> struct AirArgs {
    Grid<T> g;
    gptr<const T> U, V, W;  // mean wind (NULL when zero)
    T gam, Pav, gP, Aav;
    int I1, I2, J1, J2, K1, K2;  // active PML widths (J1 carries the y1 time gate)
    DProf<T> x1, x2, y1, y2, z1, z2;
    gptr<const T> Qo[6][4];  // slabs x1 x2 y1 y2 z1 z2, fields u v w p: before the step
    gptr<T> Qn[6][4];  // and after it
    int qn[6];                 // slab planes
    int qlo[6];                // first global plane of each slab
    int y1on;                  // y1 auxiliaries accumulate (gate open)
    // fused step only: the South face value source (0 copy of plane k=1, 1 porous plane
    // Kp-2 through the interface) and the porous pressure it reads
    int south_pore;
    gptr<const T> pore_pnn;
    int pore_nz;
    // embedded porous obstacle (FB step only): material id per node of the index box ob[],
    // NULL without obstacle; porous coefficients of the materials 1 (canopy) and 2 (trunk)
    gptr<const unsigned char> omask;
    int ob[6];
    T ogama[3], oinvb[3], oeff[3];
};

template <typename T> struct PoreArgs {
    Grid<T> g;
    T gama, invbeta, eff;
};

// All face conditions of one phase, air and porous, in one launch (lee_bc.hip).
template <typename T> struct BCArgs {
    Grid<T> air, pore;         // pore.nx == 0 without a porous ground
    int porous, iface, tdbc, y1on;
    // the x planes' interior-row nodes the air stage takes: 0 none, 1 all (the split launches,
    // rk_stage_takes_xfaces: the enumeration covers only the edge rings), 2 those of the
    // batched stage (batch_takes_face; the enumeration skips them)
    int xring;
    long long nb_air, nb_pore; // boundary nodes of each box
    gptr<const T> Qo[6][4];
    gptr<T> Qn[6][4];
    int qn[6], qlo[6], qaxis[6];
    T dt;
    gptr<const double> psi_old;  // East TDBC histories [i + nx*k][6], read
    gptr<double> psi_new;  // and written (double buffered)
    double decay[6], gain[6], inv, Zinf, tb;
    // ground impedance on the South face: histories FP64, double buffered, pole-major on the
    // device (pole q of ground node n = i + nx*j at q*nx*ny + n, so a wave's accesses to one
    // pole are contiguous; restart files and lee_get_ground_psi use node-major [n][gK]);
    // the ground interface plane k = 0: its metric (gmx[i] = 1/X_over_x, gmy[j], gmz), the
    // porous coefficients of the top material (dt/(Kp porosity), 1/rho_e, 1/(1 + beta)),
    // 2 dt/(rho_e dz) of the half-cell momentum and Zinf (1 + tb sum_k a_k)
    int ground, gK;
    gptr<const double> gpsi_old;
    gptr<double> gpsi_new;
    double gdecay[8], ggain[8], ginv, gZinf, gtb;
    gptr<const double> gmx, gmy;
    double gmz, g_gama, g_eff, g_ib, gcpl, gzi;
    // on the terrain-following grid (NULL when flat): the cross terms zX, zY of the plane's
    // inverse Jacobian per ground node c = i + nx*j, two tables of nx*ny values one after the other
    gptr<const double> gter;
    gptr<const long long> rec_idx;  // receiver samples of time level n (pressure phase)
    int nrec;
    gptr<double> rec_out;
};
// ---- RK4 (extension) ------------------------------------------------------------------
// Registers: level n (xn), the accumulator acc = xn + sum_s b_s k_s, and the destination
// (the next stage state xn + a_s k_s, or level n+1 = acc + b_4 k_4 after the last stage).
// FP32 adds a compensated accumulation: `lo` is the low part of acc within a step and `err`
// the rounding error of the stored state, carried from step to step.
template <typename T> struct RKArgs {
    AirArgs<T> a;              // stage state in g.un .. g.pn (g.pnn = g.pn), auxiliaries Qo = Qn
    gptr<const T> xn[4];  // u v w p at level n
    gptr<T> acc[4], lo[4], err[4], dst[4];
    gptr<const T> qxn[6][4];  // auxiliaries at level n, accumulator, destination
    gptr<T> qacc[6][4], qdst[6][4];
    int stage;                 // 0..3
};
template <typename T> struct RKPoreArgs {
    Grid<T> g;                 // stage state in g.un .. g.pn
    T gama, eff, beta;         // porous.cpp:44-50: dt/(Kp Omega), 1/rho_e, sigma dt/rho_e
    T cf[5];                   // drag integrating factors E, E2, E2/2, E/6, E2/3 (E = e^-beta)
    gptr<const T> xn[4];
    gptr<T> acc[4], lo[4], err[4], dst[4];
    int stage;
};
template <typename T> struct RKBndArgs {
    BCArgs<T> b;               // destination state as air/pore "nn" fields (face walks)
    gptr<const T> src[4];  // air stage state (right-hand side of the boundary auxiliaries)
    gptr<const T> qxn[6][4];
    gptr<T> qacc[6][4], qdst[6][4];
    int stage;
};
// wind: 0 none, 1 U V W fields, 2 V only (U = W = 0: velocity_method 7)
// r: the stage arguments on the host (grid shape), rd: the same in device memory (read by
// the kernel)
template <typename T>
void launch_rk_air(const RKArgs<T> &r, const RKArgs<T> *rd, int wind, bool terrain, hipStream_t s);
// Launch shapes of the RK4 air stage the autotuner chooses from (variant 0 = launch_rk_air).
struct RkTile { int ty; bool columns; const char *name; };
constexpr RkTile kRkTiles[11] = {{4, false, "64x4 planes"}, {4, true, "64x4 xcd-columns"}, {8, true, "64x8 xcd-columns"},
                                {8, false, "64x8 planes"}, {16, true, "64x16 xcd-columns"},
                                {4, false, "64x4 interior + layer shell"}, {4, false, "64x4 interior + layers by axis"},
                                {4, false, "64x4 LDS-staged interior + layer shell"},
                                {4, false, "64x4 LDS-staged interior + layers by axis"},
                                {4, false, "64x4 LDS-staged interior + LDS-staged layer shell"},
                                {4, false, "64x4 LDS-staged interior + LDS-staged layers by axis"}};
constexpr int kRkVariants = 11;
template <typename T>
void launch_rk_air_v(const RKArgs<T> &r, const RKArgs<T> *rd, int wind, bool terrain, int v, hipStream_t s);
// Whether shape v's air stage also takes the x-face nodes of the interior rows (the boundary
// stage then walks only the x planes' edge rings: BCArgs::xring).
template <typename T> bool rk_stage_takes_xfaces(const RKArgs<T> &r, int v);
// Whether the split FB passes (launch_air_p / launch_air_v) take the x-face nodes of the
// interior rows (the face passes then walk only the x planes' edge rings).
template <typename T> bool fb_passes_take_xfaces(const AirArgs<T> &a);
template <typename T> void launch_rk_pore(const RKPoreArgs<T> &r, bool terrain, hipStream_t s);
// r: the stage arguments on the host (node count), rd: the same in device memory (read by the
// kernel; by value the block was copied to scratch for the slab indexing)
template <typename T> void launch_rk_bnd(const RKBndArgs<T> &r, const RKBndArgs<T> *rd, hipStream_t s);
// Ensemble forms: args[d] in device memory for member d, first[d] = its first block
// (first[n] = nblocks); 256 boundary nodes per block for bnd.  Air and porous blocks of 256
// threads cover 256 consecutive columns i + nx*(j - 1) of the interior rows j = 1 .. ny-2 of
// one plane (the member's rows are 64-128 points, so 64-wide tiles would leave about a
// quarter of the lanes idle): batch_blocks of them per member.
constexpr int kBatchBlock = 256;
// Launch forms of the batched air and porous stages (same points, same arithmetic, identical
// bits; the ensemble's autotuner keeps the fastest):
//   0  one block per tile, member after member, plane after plane;
//   1  as 0 with each member plane's tiles padded to a multiple of 8, so that a tile and the
//      tiles above and below it run on one XCD under round-robin dispatch (padding tiles do
//      nothing);
//   2  persistent: as many blocks as the chip holds at once, each walking the tiles of form 0
//      with a stride of the grid size.
constexpr int kEnsForms = 3;
inline long long batch_blocks(int nx, int ny, int nz, bool pad = false)
{
    long long per = ((long long)nx * (ny - 2) + kBatchBlock - 1) / kBatchBlock;
    if (pad) per = (per + 7) & ~7LL;
    return per * (nz - 2);
}
// Whether node (i, j) of an interior row is an x-face node the batched air stage takes: its
// interior neighbour (i + 1 or i - 1, the next or previous column) is in the same wavefront
// of 64 consecutive columns (blocks start at multiples of 256 columns).  The boundary stage
// skips exactly these (BCArgs::xring 2).
__host__ __device__ inline bool batch_takes_face(int nx, int i, int j)
{
    if (i != 0 && i != nx - 1) return false;
    const int c = (j - 1) * nx + i;
    return i == 0 ? (c & 63) != 63 : (c & 63) != 0;
}
// form: kEnsForms; first and nblocks count the tiles of that form (padded for form 1)
template <typename T>
void launch_rk_air_batch(const RKArgs<T> *args, const int *first, int n, int nblocks, int wind, bool terrain, int form,
                         hipStream_t s);
template <typename T>
void launch_rk_pore_batch(const RKPoreArgs<T> *args, const int *first, int n, int nblocks, bool terrain, int form,
                          hipStream_t s);
template <typename T> void launch_rk_bnd_batch(const RKBndArgs<T> *args, const int *first, int n, int nblocks, hipStream_t s);
template <typename T> void launch_sample_batch(const T *const *p, const long long *idx, int n, double *out, hipStream_t s);

template <typename T> long long boundary_count(int nx, int ny, int nz, bool xring = false);
template <typename T> void launch_bc_p_all(const BCArgs<T> &a, hipStream_t s);
template <typename T> void launch_bc_v_all(const BCArgs<T> &a, hipStream_t s);

// pad: the box lists of the split launches with each plane's tiles padded to a multiple of 8
// (a tile column on one XCD under round-robin dispatch; padding tiles do nothing).
// vec_ok: the interior box may run the vector form (several points per lane, FP32 only)
template <typename T> void launch_air_p(const AirArgs<T> &a, bool wind, hipStream_t s, bool pad = false,
                                          bool vec_ok = true);
template <typename T> void launch_air_v(const AirArgs<T> &a, bool wind, hipStream_t s, bool pad = false,
                                          bool vec_ok = true);
// Both air phases in one pass (pressure then velocity, plane by plane); see lee_kernels.hip.
// Returns false when the grid does not fit the fused kernel (nx - 2 > 1024).
template <typename T> bool fused_fits(int nx);
// per_cu: the pencil depth is the planes / chunks that give about per_cu blocks per CU
template <typename T> void launch_air_fb(const AirArgs<T> &a, bool wind, hipStream_t s, int per_cu = 4);
// Launch forms of the FB_p_v air passes (the FB autotuner picks one; identical bits): 0 the
// split passes, 1 the split passes with padded box lists, 2 .. 5 the fused pass with pencils
// for kFbPerCu[form - 2] blocks per CU, 6 the split passes with the one-point interior.
constexpr int kFbPerCu[4] = {2, 4, 8, 16};
constexpr int kFbForms = 7;
constexpr bool fb_form_fused(int f) { return f >= 2 && f < 6; }
constexpr unsigned kFbSplitForms = 1u | 2u | 64u, kFbFusedForms = 4u | 8u | 16u | 32u;
template <typename T> void launch_pore_p(const PoreArgs<T> &a, hipStream_t s);
template <typename T> void launch_pore_v(const PoreArgs<T> &a, hipStream_t s);
// Moving-frame shift along j of an (A x B x C) box (B = frame length), out of place.
template <typename T> void launch_shift(T *dst, const T *src, int A, int B, int C, int tempJ, int lead,
                                        int s, hipStream_t st);
template <typename T>
void launch_shift_yslab(T *dst, const T *src, int nx, int n, int nz, int lo, int ny, int tempJ, int lead, int sh,
                        hipStream_t st);
template <typename T> void launch_copy_plane_j(T *p, int nx, int ny, int nz, hipStream_t s);
template <typename T> void launch_sample(const T *p, const long long *idx, int n, double *out, hipStream_t s);
void launch_copy_words(float *dst, const float *src, size_t n, hipStream_t s);
// Many-stream reference kernel: reads nr and writes nw arrays of n words, one word of each per
// lane; (nr, nw) one of (1,1) (4,4) (8,4) (12,8) (20,8) (28,8), else nothing is launched and
// the result is false.
constexpr int kStreamMaxR = 28, kStreamMaxW = 8;
bool launch_stream_words(float *const *out, int nw, const float *const *in, int nr, size_t n, hipStream_t s);
template <typename T> void launch_finite_check(const T *f, size_t n, int *flag, hipStream_t s);
template <typename T>
void launch_wind_log(T *V, int nx, int ny, int nz, const double *h, const double *az, const double *tk, double ustar,
                     double z0, hipStream_t s);
template <typename T>
void launch_add_source(T *p, int nx, int ny, int nz, const double *X, const double *Y, const double *Z, double sx,
                       double sy, double sz, double gc, double amp, hipStream_t s);
template <typename T> void launch_subsample(const T *p, int nx, int ny, int nz, int stride, int ox, int oy, int oz,
                                             double *out, hipStream_t s);

}  // namespace lee
