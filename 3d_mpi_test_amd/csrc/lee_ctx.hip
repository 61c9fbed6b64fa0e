// lee_ctx.hip -- solver context and the C ABI of include/lee_abi.h.
//
// One context = one domain in HBM on one device, one HIP stream.  The host keeps the
// FP64 model (lee_model.cpp) and drives the kernels of lee_kernels.hip in the phase order
// of airpore::Cal_pressure / Cal_velocity (airpore.cpp:324-398).  Time levels n and n+1
// are two buffer sets whose roles swap after every step instead of the reference's full
// nn -> n copy (calculation_2D.cpp:354-369): every point of level n+1 is rewritten by the
// interior update or by a face condition before it is read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lee_abi.h"
#include "lee_kernels.h"
#include "lee_model.h"

namespace lee {
thread_local std::string g_global_err;
void set_global_error(const std::string &m) { g_global_err = m; }
}  // namespace lee

using namespace lee;

namespace {

constexpr int kRecChunk = 4096;  // receiver samples buffered on the device between copies
constexpr int kGraphSteps = 64;  // steps per captured graph (use_graph)

struct DevArr {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct lee_ctx {
    Model m;
    int prec = 64;
    size_t esz = 8;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::vector<DevArr> allocs;
    // fields: [0..3] level A (u v w p), [4..7] level B; `cur` selects which is level n
    void *air[8] = {};
    void *pore[8] = {};
    int cur = 0;
    void *U = nullptr, *V = nullptr, *W = nullptr;
    void *mx = nullptr, *my = nullptr, *mz = nullptr, *pmx = nullptr, *pmy = nullptr, *pmz = nullptr;
    void *prof[6][3] = {};
    void *Q[2][6][4] = {};  // PML auxiliaries, two sets: Q[qcur] holds the values before a step
    int qcur = 0;
    bool fused = false;     // both air phases in one kernel (k_air_fb)
    int fb_form = -1;       // FB_p_v air-pass launch form (kFbForms), -1 until tuned
    unsigned fb_mask = 0;   // the forms lee_options.path allows, bit f for form f
    int rk_variant = -1;    // RK4 air-stage launch shape (kRkTiles index), -1 until tuned
    bool rk_xfaces = false; // that shape's air stage takes the x-face nodes of the interior rows
    bool fb_xfaces = false; // the split FB passes take them
    // RK4 air-stage arguments of the four stages in device memory and the host copy they were
    // uploaded from (refreshed when buffer roles or the y1 gate change, never inside a capture)
    void *rk_args_dev = nullptr;
    std::vector<unsigned char> rk_args_host, rk_args_key;
    bool capturing = false;
    bool args_fault = false;  // rk_args_sync failed inside an enqueue (reported by run_steps)
    // RK4 registers: second stage state, accumulator, FP32 accumulator low part and state
    // error ([0] air, [1] porous; u v w p; the air accumulator is one block of the four
    // fields interleaved, at rk_acc[0][0]), and of the auxiliaries (stage state, accumulator)
    void *rk_sb[2][4] = {}, *rk_acc[2][4] = {}, *rk_lo[2][4] = {}, *rk_err[2][4] = {};
    void *rkq_sb[6][4] = {};
    void *rkq_acc[6][4] = {};  // [s][0]: the slab's four accumulators interleaved ([s][1..3] unused)
    // terrain tables (sh, hx, hy per (i,j); tk per k) of the air [0] and porous [1] grids
    void *ter[2][4] = {};
    unsigned char *omask = nullptr;  // embedded obstacle material ids (FB)
    // log wind on the device: terrain height of the frame [nx*ny], z and tk per level [2*nz]
    double *wind_h = nullptr, *wind_z = nullptr;
    bool wind_flat_done = false;
    // time-series source: amplitudes, period in steps, node coordinates X | Y(frame) | Z
    std::vector<double> src_ps;
    int src_period = 0;
    double *src_xyz = nullptr;
    // captured step chunks (use_graph): keyed by the buffer roles they start from
    struct GraphEntry { int n, cur, qcur, gate_on; void *psi, *gpsi; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs;
    double *psi = nullptr, *psi2 = nullptr;  // TDBC histories, read / written each step
    double *gpsi = nullptr, *gpsi2 = nullptr;  // ground impedance histories (South face)
    double *gmx = nullptr, *gmy = nullptr;     // ground plane metric 1/X_over_x, 1/Y_over_y
    double gmz = 0;                            // and 1/Z_over_z of the plane k = 1
    // over terrain, the cross terms zX, zY of the plane's inverse Jacobian per ground node,
    // two tables of nx*ny values one after the other
    double *gter = nullptr;
    long long *rec_idx = nullptr;
    double *rec_buf = nullptr;
    int *flag = nullptr;
    void *shift_tmp = nullptr;  // frame shift: one field or slab copied out of place
    size_t shift_tmp_bytes = 0;
    int device = 0;
    int mf_count = -1;
    long long steps = 0, counter = 0;
    float last_ms = 0.f;
    std::vector<double> host_rec;
    // profiling: event pairs around the stencil launches, resolved at the end of lee_step
    int profile = 0;  // lee_profile: 0 off, 1 every kernel kind, 2 the air stencil launches only
    std::vector<hipEvent_t> evpool;
    size_t evnext = 0;
    struct Pending { int kind; hipEvent_t a, b; };
    std::vector<Pending> pending;
    lee_kstat kstat[10] = {};

    ~lee_ctx()
    {
        // teardown: errors cannot be reported from a destructor
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        for (auto &a : allocs) if (a.p) (void)hipFree(a.p);
        for (auto e : evpool) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int set_err(lee_ctx *c, int code, const std::string &m)
{
    if (c) c->err = m;
    set_global_error(m);
    return code;
}

#define HIPCHK(c, expr)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return set_err((c), LEE_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// ---- device staging of lee_write_contour_noalloc ----------------------------------------
// A file-scope __device__ array exists once on every GPU the library runs on, from the loading
// of this code object until the process ends, so that entry point stages through it without
// allocating.  On the host the array's name is a stand-in, not a device address: host code
// never passes it to a kernel or a copy.  lee_create looks up each device's address once, with
// that device current (hipGetSymbolAddress), and the entry point uses that address.  Calls on
// one GPU share its array: two such calls that overlap on one GPU race.
constexpr size_t kScratchBytes = (size_t)4 << 20;  // 4 MiB per GPU
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_at[kMaxDevices] = {};  // each device's address of `scratch`

// Once per device, with that device current.
int scratch_setup(lee_ctx *c, int device)
{
    if (device < 0 || device >= kMaxDevices) return set_err(c, LEE_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_at[device]) return LEE_OK;
    void *p = nullptr;
    HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(c, LEE_ERR_DEVICE, "no device address for the contour staging array");
    g_scratch_at[device] = (double *)p;
    return LEE_OK;
}

double *scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_at[device] : nullptr;
}

int dalloc(lee_ctx *c, void **out, size_t bytes)
{
    *out = nullptr;
    if (bytes == 0) return LEE_OK;
    void *p = nullptr;
    HIPCHK(c, hipMalloc(&p, bytes));
    HIPCHK(c, hipMemsetAsync(p, 0, bytes, c->stream));
    c->allocs.push_back({p, bytes});
    *out = p;
    return LEE_OK;
}

// Upload FP64 host values as the device type.
int upload(lee_ctx *c, void *dst, const double *src, size_t n)
{
    if (!n) return LEE_OK;
    // set-up path only: wait for the copy so the host buffer may be released on return
    if (c->prec == 64) {
        HIPCHK(c, hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(src, src + n);
        HIPCHK(c, hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return LEE_OK;
}

// Upload FP64 host values as FP64 whatever the state precision (coordinates).
int upload_f64(lee_ctx *c, double *dst, const double *src, size_t n)
{
    if (!n) return LEE_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LEE_OK;
}

int download(lee_ctx *c, double *dst, const void *src, size_t n)
{
    if (!n || !dst) return LEE_OK;
    if (c->prec == 64) {
        HIPCHK(c, hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < n; ++i) dst[i] = tmp[i];
    }
    return LEE_OK;
}

enum { KS_AIR_P = 0, KS_AIR_V = 1, KS_PORE_P = 2, KS_PORE_V = 3, KS_AIR_FB = 4, KS_BC_P = 5, KS_BC_V = 6,
       KS_RK_AIR = 7, KS_RK_PORE = 8, KS_RK_BND = 9, KS_N = 10 };

hipEvent_t pooled_event(lee_ctx *c)
{
    if (c->evnext == c->evpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->evpool.push_back(e);
    }
    return c->evpool[c->evnext++];
}

// Record an event before (begin=true) or after a timed launch.
void prof_mark(lee_ctx *c, int kind, bool begin)
{
    if (!c->profile) return;
    if (c->profile == 2 && kind != KS_RK_AIR && kind != KS_AIR_P && kind != KS_AIR_V && kind != KS_AIR_FB) return;
    hipEvent_t e = pooled_event(c);
    if (!e) return;
    (void)hipEventRecord(e, c->stream);
    if (begin) c->pending.push_back({kind, e, nullptr});
    else c->pending.back().b = e;
}

void prof_resolve(lee_ctx *c)
{
    for (auto &p : c->pending) {
        float ms = 0.f;
        if (p.a && p.b && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->kstat[p.kind].launches += 1;
            c->kstat[p.kind].total_ms += ms;
        }
    }
    c->pending.clear();
    c->evnext = 0;
}

size_t air_n(const lee_ctx *c) { return (size_t)c->m.nx * c->m.ny * c->m.nz; }
size_t pore_n(const lee_ctx *c) { return (size_t)c->m.pnx * c->m.pny * c->m.pnz; }

template <typename T> Grid<T> air_grid(lee_ctx *c)
{
    Grid<T> g;
    const Model &m = c->m;
    int n = c->cur * 4, o = 4 - n;
    g.nx = m.nx; g.ny = m.ny; g.nz = m.nz;
    g.un = (T *)c->air[n + 0]; g.vn = (T *)c->air[n + 1]; g.wn = (T *)c->air[n + 2]; g.pn = (T *)c->air[n + 3];
    g.unn = (T *)c->air[o + 0]; g.vnn = (T *)c->air[o + 1]; g.wnn = (T *)c->air[o + 2]; g.pnn = (T *)c->air[o + 3];
    g.mx = (const T *)c->mx; g.my = (const T *)c->my; g.mz = (const T *)c->mz;
    g.dxi = (T)m.dxi; g.dyi = (T)m.dyi; g.dzi = (T)m.dzi; g.dt = (T)m.dt;
    g.tsh = (const T *)c->ter[0][0]; g.thx = (const T *)c->ter[0][1]; g.thy = (const T *)c->ter[0][2]; g.ttk = (const T *)c->ter[0][3];
    return g;
}

template <typename T> Grid<T> pore_grid(lee_ctx *c)
{
    Grid<T> g;
    const Model &m = c->m;
    int n = c->cur * 4, o = 4 - n;
    g.nx = m.pnx; g.ny = m.pny; g.nz = m.pnz;
    g.un = (T *)c->pore[n + 0]; g.vn = (T *)c->pore[n + 1]; g.wn = (T *)c->pore[n + 2]; g.pn = (T *)c->pore[n + 3];
    g.unn = (T *)c->pore[o + 0]; g.vnn = (T *)c->pore[o + 1]; g.wnn = (T *)c->pore[o + 2]; g.pnn = (T *)c->pore[o + 3];
    g.mx = (const T *)c->pmx; g.my = (const T *)c->pmy; g.mz = (const T *)c->pmz;
    g.dxi = (T)m.pdxi; g.dyi = (T)m.pdyi; g.dzi = (T)m.pdzi; g.dt = (T)m.dt;
    g.tsh = (const T *)c->ter[1][0]; g.thx = (const T *)c->ter[1][1]; g.thy = (const T *)c->ter[1][2]; g.ttk = (const T *)c->ter[1][3];
    return g;
}

template <typename T> AirArgs<T> air_args(lee_ctx *c, int J1, int y1on)
{
    const Model &m = c->m;
    AirArgs<T> a;
    a.g = air_grid<T>(c);
    a.U = (const T *)c->U; a.V = (const T *)c->V; a.W = (const T *)c->W;
    a.gam = (T)m.gamma; a.Pav = (T)m.Pav; a.gP = (T)m.gP; a.Aav = (T)m.Aav;
    a.I1 = m.I1; a.I2 = m.I2; a.J1 = J1; a.J2 = m.J2; a.K1 = m.K1; a.K2 = m.K2;
    DProf<T> *pr[6] = {&a.x1, &a.x2, &a.y1, &a.y2, &a.z1, &a.z2};
    for (int s = 0; s < 6; ++s) {
        pr[s]->s = (const T *)c->prof[s][0]; pr[s]->c = (const T *)c->prof[s][1]; pr[s]->d = (const T *)c->prof[s][2];
        for (int f = 0; f < 4; ++f) {
            a.Qo[s][f] = (const T *)c->Q[c->qcur][s][f];
            a.Qn[s][f] = (T *)c->Q[c->qcur ^ 1][s][f];
        }
        a.qn[s] = m.slab[s].n;
        a.qlo[s] = m.slab[s].lo;
    }
    a.y1on = y1on;
    a.south_pore = m.south_porous && m.iface_copy;
    a.omask = c->omask;
    for (int q = 0; q < 6; ++q) a.ob[q] = m.ob[q];
    for (int q = 0; q < 3; ++q) { a.ogama[q] = (T)m.o_gama[q]; a.oinvb[q] = (T)m.o_invbeta[q]; a.oeff[q] = (T)m.o_eff[q]; }
    a.pore_pnn = m.south_porous ? pore_grid<T>(c).pnn : nullptr;
    a.pore_nz = m.pnz;
    return a;
}

template <typename T> BCArgs<T> bc_args(lee_ctx *c, const AirArgs<T> &aa, double *rec_out)
{
    const Model &m = c->m;
    BCArgs<T> b;
    b.air = aa.g;
    if (m.south_porous) b.pore = pore_grid<T>(c);
    else { std::memset(&b.pore, 0, sizeof b.pore); }
    b.porous = m.south_porous; b.iface = m.south_porous && m.iface_copy; b.tdbc = m.east_tdbc; b.y1on = aa.y1on;
    b.xring = 0;  // the boundary passes walk whole x planes unless a launch shape takes their rows
    b.nb_air = boundary_count<T>(m.nx, m.ny, m.nz);
    b.nb_pore = m.south_porous ? boundary_count<T>(m.pnx, m.pny, m.pnz) : 0;
    for (int s = 0; s < 6; ++s) {
        for (int f = 0; f < 4; ++f) { b.Qo[s][f] = aa.Qo[s][f]; b.Qn[s][f] = aa.Qn[s][f]; }
        b.qn[s] = m.slab[s].n; b.qlo[s] = m.slab[s].lo; b.qaxis[s] = m.slab[s].axis;
    }
    b.dt = (T)m.dt;
    b.psi_old = c->psi; b.psi_new = c->psi2;
    for (int q = 0; q < 6; ++q) { b.decay[q] = m.td_decay[q]; b.gain[q] = m.td_gain[q]; }
    b.inv = m.td_inv; b.Zinf = m.Zinf; b.tb = m.dt / m.tau;
    b.ground = m.ground; b.gK = m.gK; b.gpsi_old = c->gpsi; b.gpsi_new = c->gpsi2;
    for (int q = 0; q < 8; ++q) { b.gdecay[q] = m.g_decay[q]; b.ggain[q] = m.g_gain[q]; }
    b.ginv = m.g_inv; b.gZinf = m.gZinf; b.gtb = m.g_tb;
    b.gmx = c->gmx; b.gmy = c->gmy; b.gmz = c->gmz;
    b.gter = c->gter;
    b.g_gama = m.g_gama; b.g_eff = m.g_eff; b.g_ib = m.g_ib;
    b.gcpl = 2.0 * m.dt * m.g_eff * c->gmz * m.dzi;
    b.gzi = m.ground ? m.gZinf / m.g_inv : 0.0;
    b.rec_idx = c->rec_idx; b.nrec = (int)m.ri.size(); b.rec_out = rec_out;
    return b;
}

// Upload the per-frame tables (metrics, PML profiles, wind) of frame N.
int upload_frame(lee_ctx *c, int N)
{
    const Model &m = c->m;
    FrameTables t;
    m.frame_tables(N, t, false);
    auto metric = [&](void *dst, const std::vector<double> &v) {
        // FP32 on the separable grid stores the reciprocals (the metric itself); the terrain
        // kernels form the Jacobian per point from X_over_x ... in either precision
        if (c->prec == 64 || m.terrain) return upload(c, dst, v.data(), v.size());
        std::vector<double> r(v.size(), 0.0);
        for (size_t i = 1; i + 1 < v.size(); ++i) r[i] = 1.0 / v[i];  // separable metric in FP64
        return upload(c, dst, r.data(), r.size());
    };
    int rc;
    if ((rc = metric(c->mx, t.Xx)) || (rc = metric(c->my, t.Yy)) || (rc = metric(c->mz, t.Zz))) return rc;
    if (m.south_porous)
        if ((rc = metric(c->pmx, t.pXx)) || (rc = metric(c->pmy, t.pYy)) || (rc = metric(c->pmz, t.pZz))) return rc;
    if (m.ground) {
        std::vector<double> gx(m.nx, 0.0), gy(m.ny, 0.0);
        for (int i = 1; i + 1 < m.nx; ++i) gx[i] = 1.0 / t.Xx[i];
        for (int j = 1; j + 1 < m.ny; ++j) gy[j] = 1.0 / t.Yy[j];
        c->gmz = 1.0 / t.Zz[1];
        // FP64 in both precisions (the plane's coefficients are cast per use)
        HIPCHK(c, hipMemcpyAsync(c->gmx, gx.data(), gx.size() * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->gmy, gy.data(), gy.size() * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (m.terrain) {
            // the plane's metric on the terrain-following grid: that of a porous top plane
            // shifted rigidly with the ground (Z_over_z = Zz, Z_over_x = hx, Z_over_y = hy)
            const size_t n2 = (size_t)m.nx * m.ny;
            std::vector<double> tab(2 * n2, 0.0);
            double *zx = tab.data(), *zy = zx + n2;
            for (int j = 1; j + 1 < m.ny; ++j)
                for (int i = 1; i + 1 < m.nx; ++i) {
                    const size_t q = (size_t)i + (size_t)m.nx * j;
                    zx[q] = -(t.ter.hx[q] * gx[i]) * c->gmz;
                    zy[q] = -(t.ter.hy[q] * gy[j]) * c->gmz;
                }
            if ((rc = upload_f64(c, c->gter, tab.data(), tab.size()))) return rc;
        }
    }
    const Profile *P[6] = {&t.x1, &t.x2, &t.y1, &t.y2, &t.z1, &t.z2};
    for (int s = 0; s < 6; ++s) {
        if ((rc = upload(c, c->prof[s][0], P[s]->s.data(), P[s]->s.size()))) return rc;
        if ((rc = upload(c, c->prof[s][1], P[s]->c.data(), P[s]->c.size()))) return rc;
        if ((rc = upload(c, c->prof[s][2], P[s]->d.data(), P[s]->d.size()))) return rc;
    }
    if (m.has_wind && m.wind_v_only) {
        // velocity_method 7 on the device: over terrain the ground under a node changes with
        // the frame; on a flat grid the field is the same in every frame
        const bool ter = m.terrain;
        if (ter || !c->wind_flat_done) {
            if (ter && (rc = upload_f64(c, c->wind_h, t.ter.h.data(), t.ter.h.size()))) return rc;
            if (c->prec == 64)
                launch_wind_log<double>((double *)c->V, m.nx, m.ny, m.nz, ter ? c->wind_h : nullptr, c->wind_z,
                                        c->wind_z + m.nz, m.opt.wind_ustar, m.opt.wind_z0, c->stream);
            else
                launch_wind_log<float>((float *)c->V, m.nx, m.ny, m.nz, ter ? c->wind_h : nullptr, c->wind_z,
                                       c->wind_z + m.nz, m.opt.wind_ustar, m.opt.wind_z0, c->stream);
            HIPCHK(c, hipGetLastError());
            c->wind_flat_done = true;
        }
    } else if (m.has_wind) {
        if ((rc = upload(c, c->U, t.U.data(), t.U.size())) || (rc = upload(c, c->V, t.V.data(), t.V.size())) ||
            (rc = upload(c, c->W, t.W.data(), t.W.size())))
            return rc;
    }
    if (m.terrain) {
        const TerrainTables *tt[2] = {&t.ter, &t.pter};
        for (int d = 0; d < (m.south_porous ? 2 : 1); ++d) {
            const std::vector<double> *v[4] = {&tt[d]->sh, &tt[d]->hx, &tt[d]->hy, &tt[d]->tk};
            for (int q = 0; q < 4; ++q)
                if ((rc = upload(c, c->ter[d][q], v[q]->data(), v[q]->size()))) return rc;
        }
    }
    if (c->src_xyz)
        if ((rc = upload_f64(c, c->src_xyz + m.nx, t.Y.data(), t.Y.size()))) return rc;
    // receiver linear indices in this frame
    int nr = (int)m.ri.size();
    std::vector<long long> idx(nr);
    for (int r = 0; r < nr; ++r) idx[r] = m.receiver_index(r, N);
    if (nr) HIPCHK(c, hipMemcpyAsync(c->rec_idx, idx.data(), nr * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LEE_OK;
}

template <typename T> int frame_shift(lee_ctx *c, int N)
{
    const Model &m = c->m;
    int tempJ = m.d.mf_jmax - m.lead;  // calculation_2D.cpp:459 (temp_J of frame N > 0)
    if (N == 0 || tempJ == 0) return LEE_OK;
    int s = m.trail - 1;
    int n = c->cur * 4, o = 4 - n;
    // fields: level n -> (free) level n+1 buffers, then swap roles
    for (int f = 0; f < 4; ++f)
        launch_shift<T>((T *)c->air[o + f], (const T *)c->air[n + f], m.nx, m.ny, m.nz, tempJ, m.lead, s, c->stream);
    if (m.south_porous)
        for (int f = 0; f < 4; ++f)
            launch_shift<T>((T *)c->pore[o + f], (const T *)c->pore[n + f], m.pnx, m.pny, m.pnz, tempJ, m.lead, s, c->stream);
    c->cur ^= 1;
    n = c->cur * 4;
    // the FP32 RK4 state error moves with the state (through the second stage buffer)
    if (m.rk4 && c->rk_err[0][0])
        for (int d = 0; d < (m.south_porous ? 2 : 1); ++d)
            for (int f = 0; f < 4; ++f) {
                const int A = d ? m.pnx : m.nx, B = d ? m.pny : m.ny, C = d ? m.pnz : m.nz;
                launch_shift<T>((T *)c->rk_sb[d][f], (const T *)c->rk_err[d][f], A, B, C, tempJ, m.lead, s, c->stream);
                std::swap(c->rk_sb[d][f], c->rk_err[d][f]);
            }
    launch_copy_plane_j<T>((T *)c->air[n + 3], m.nx, m.ny, m.nz, c->stream);
    if (m.south_porous) launch_copy_plane_j<T>((T *)c->pore[n + 3], m.pnx, m.pny, m.pnz, c->stream);
    // PML auxiliaries of the x and z layers (air.cpp:2522-2729); y layers stay
    for (int sl : {0, 1, 4, 5}) {
        const SlabGeom &g = m.slab[sl];
        if (!g.n) continue;
        int A = g.axis == 0 ? g.n : m.nx, C = g.axis == 2 ? g.n : m.nz;
        size_t cnt = (size_t)A * m.ny * C;
        for (int f = 0; f < 4; ++f) {
            void *q = c->Q[c->qcur][sl][f];
            HIPCHK(c, hipMemcpyAsync(c->shift_tmp, q, cnt * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
            launch_shift<T>((T *)q, (const T *)c->shift_tmp, A, m.ny, C, tempJ, m.lead, s, c->stream);
        }
    }
    // RK4: the y layers move with the frame as well (their auxiliaries meet the shifted x and
    // z layers' in the corner cross terms; the reference's FB step leaves them in place)
    if (m.rk4)
        for (int sl : {2, 3}) {
            const SlabGeom &g = m.slab[sl];
            if (!g.n) continue;
            size_t cnt = (size_t)m.nx * g.n * m.nz;
            const int lo = g.lo;
            for (int f = 0; f < 4; ++f) {
                void *q = c->Q[c->qcur][sl][f];
                HIPCHK(c, hipMemcpyAsync(c->shift_tmp, q, cnt * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
                launch_shift_yslab<T>((T *)q, (const T *)c->shift_tmp, m.nx, g.n, m.nz, lo, m.ny, tempJ, m.lead, s, c->stream);
            }
        }
    HIPCHK(c, hipGetLastError());
    return LEE_OK;
}

// ---- RK4 (extension): four stages per step; level n stays in the current level buffers
// (frame shifts, state access and receivers work as for FB), the stage states alternate
// between the other level buffers (A) and rk_sb (B):  X -> A -> B -> A -> X.
// Ground impedance histories: pole-major on the device (Psi_q of node n at q * nn + n, so a
// wave's accesses to one pole are contiguous), node-major in restart files and in
// lee_get_ground_psi (n * K + q).  to_nodes: device order -> node-major, else the reverse.
static void ground_psi_layout(const double *src, double *dst, size_t nn, int K, bool to_nodes)
{
    for (size_t n = 0; n < nn; ++n)
        for (int q = 0; q < K; ++q) {
            if (to_nodes) dst[n * K + q] = src[(size_t)q * nn + n];
            else dst[(size_t)q * nn + n] = src[n * K + q];
        }
}

template <typename T> void rk_stage_buffers(lee_ctx *c, int d, int st, T *src[4], T *dst[4])
{
    void **lev = d ? c->pore : c->air;
    T *X[4], *A[4], *B[4];
    for (int f = 0; f < 4; ++f) {
        X[f] = (T *)lev[c->cur * 4 + f];
        A[f] = (T *)lev[(c->cur ^ 1) * 4 + f];
        B[f] = (T *)c->rk_sb[d][f];
    }
    T **S[4] = {X, A, B, A}, **D[4] = {A, B, A, X};
    for (int f = 0; f < 4; ++f) { src[f] = S[st][f]; dst[f] = D[st][f]; }
}

template <typename T> void rk_stage_q(lee_ctx *c, int st, int s, int f, T *&src, T *&dst)
{
    T *X = (T *)c->Q[c->qcur][s][f], *A = (T *)c->Q[c->qcur ^ 1][s][f], *B = (T *)c->rkq_sb[s][f];
    T *S[4] = {X, A, B, A}, *D[4] = {A, B, A, X};
    src = S[st];
    dst = D[st];
}

// y1 time gate of the coming step (air.cpp:1965-1969): advances the counter.
void gate_tick(lee_ctx *c, int &J1, int &y1on)
{
    const Model &m = c->m;
    const bool gate = (m.judge[1] == 1 || m.judge[1] == 3);
    J1 = m.J1;
    if (gate) {
        c->counter += 1;
        J1 = c->counter >= m.pml_time ? m.J3 : 0;
    }
    y1on = (m.slab[2].n > 0 && gate && c->counter >= m.pml_time) ? 1 : 0;
}

// calculation_2D::UpdateSource (calculation_2D.cpp:371-385) after step N (counted from 1,
// airpore.cpp:572): every `period` steps the pressure of level n+1 gains ps[N/period] times
// the initial Gaussian about the deck's source (the reference's call, airpore.cpp:608-611).
template <typename T> void inject_source(lee_ctx *c, T *p)
{
    const long long N = c->steps + 1;
    if (c->src_ps.empty() || N % c->src_period != 0 || N / c->src_period >= (long long)c->src_ps.size()) return;
    const Model &m = c->m;
    const double gc = 4.0 * std::log(2.0) / std::pow(m.d.gauss_width, 2);  // calculation_2D.cpp:26
    launch_add_source<T>(p, m.nx, m.ny, m.nz, c->src_xyz, c->src_xyz + m.nx, c->src_xyz + m.nx + m.ny, m.d.src[0],
                         m.d.src[1], m.d.src[2], gc, c->src_ps[N / c->src_period], c->stream);
}

// Kernel arguments of RK4 stage `st` for the y1 gate state (J1, y1on): stage state -> next
// stage state of the air, of the porous layer, and the face walks of the destination.
template <typename T>
void rk4_stage_args(lee_ctx *c, int st, int J1, int y1on, RKArgs<T> &r, RKPoreArgs<T> &pr, RKBndArgs<T> &b)
{
    const Model &m = c->m;
    T *src[4], *dst[4];
    rk_stage_buffers<T>(c, 0, st, src, dst);
    r = RKArgs<T>{};
    r.a = air_args<T>(c, J1, y1on);
    Grid<T> &g = r.a.g;
    g.un = src[0]; g.vn = src[1]; g.wn = src[2]; g.pn = src[3];
    g.unn = g.vnn = g.wnn = nullptr;
    g.pnn = src[3];  // the velocity increment reads the stage pressure
    for (int f = 0; f < 4; ++f) {
        r.xn[f] = (const T *)c->air[c->cur * 4 + f];
        r.acc[f] = (T *)c->rk_acc[0][0];  // one block of four interleaved fields
        r.lo[f] = (T *)c->rk_lo[0][f]; r.err[f] = (T *)c->rk_err[0][f];
        r.dst[f] = dst[f];
    }
    for (int q = 0; q < 6; ++q)
        for (int f = 0; f < 4; ++f) {
            T *qs, *qd;
            rk_stage_q<T>(c, st, q, f, qs, qd);
            r.a.Qo[q][f] = qs; r.a.Qn[q][f] = qs;
            r.qxn[q][f] = (const T *)c->Q[c->qcur][q][f];
            r.qacc[q][f] = (T *)c->rkq_acc[q][0];  // one block of four interleaved fields
            r.qdst[q][f] = qd;
        }
    r.stage = st;
    T *psrc[4] = {}, *pdst[4] = {};
    pr = RKPoreArgs<T>{};
    if (m.south_porous) {
        rk_stage_buffers<T>(c, 1, st, psrc, pdst);
        pr.g = pore_grid<T>(c);
        pr.g.un = psrc[0]; pr.g.vn = psrc[1]; pr.g.wn = psrc[2]; pr.g.pn = psrc[3];
        pr.g.unn = pr.g.vnn = pr.g.wnn = pr.g.pnn = nullptr;
        pr.gama = (T)m.p_gama; pr.eff = (T)m.p_eff; pr.beta = (T)m.p_beta;
        const double E = std::exp(-m.p_beta), E2 = std::exp(-0.5 * m.p_beta);
        const double cf[5] = {E, E2, E2 * 0.5, E / 6.0, E2 / 3.0};
        for (int q = 0; q < 5; ++q) pr.cf[q] = (T)cf[q];
        for (int f = 0; f < 4; ++f) {
            pr.xn[f] = (const T *)c->pore[c->cur * 4 + f];
            pr.acc[f] = (T *)c->rk_acc[1][f]; pr.lo[f] = (T *)c->rk_lo[1][f]; pr.err[f] = (T *)c->rk_err[1][f];
            pr.dst[f] = pdst[f];
        }
        pr.stage = st;
    }
    // faces of the destination state and the boundary-node auxiliaries
    b = RKBndArgs<T>{};
    b.b = bc_args<T>(c, r.a, nullptr);
    b.b.tdbc = 0;
    b.b.nrec = 0;
    c->rk_xfaces = rk_stage_takes_xfaces<T>(r, c->rk_variant);
    if (c->rk_xfaces) {
        // the air stage's launch shape takes the x-face nodes of the interior rows
        b.b.xring = 1;
        b.b.nb_air = boundary_count<T>(m.nx, m.ny, m.nz, true);
    }
    b.b.air.unn = dst[0]; b.b.air.vnn = dst[1]; b.b.air.wnn = dst[2]; b.b.air.pnn = dst[3];
    if (m.south_porous) {
        b.b.pore.unn = pdst[0]; b.b.pore.vnn = pdst[1]; b.b.pore.wnn = pdst[2]; b.b.pore.pnn = pdst[3];
    }
    for (int f = 0; f < 4; ++f) b.src[f] = src[f];
    for (int q = 0; q < 6; ++q)
        for (int f = 0; f < 4; ++f) { b.qxn[q][f] = r.qxn[q][f]; b.qacc[q][f] = r.qacc[q][f]; b.qdst[q][f] = r.qdst[q][f]; }
    b.stage = st;
}

int rk_wind_mode(const Model &m) { return m.has_wind ? (m.wind_v_only ? 2 : 1) : 0; }

// Brings the device copy of the four stage arguments (air stage, then boundary stage: four
// RKArgs followed by four RKBndArgs) up to date for the gate state (J1, y1on).  A change is rare (frame shifts, the y1 gate opening) and waits for the copy.
// The arguments depend on the buffer roles (level and auxiliary sets, the FP32 error and
// second stage buffers a frame shift swaps), the gate state and the air stage's launch shape
// (which may take the x-face nodes from the boundary stage): that is the key.
template <typename T> int rk_args_sync(lee_ctx *c, int J1, int y1on)
{
    const size_t sz = sizeof(RKArgs<T>), szb = sizeof(RKBndArgs<T>);
    static_assert(sizeof(RKArgs<T>) % alignof(RKBndArgs<T>) == 0, "boundary blocks follow the air blocks");
    const uintptr_t key[9] = {(uintptr_t)c->cur, (uintptr_t)c->qcur, (uintptr_t)J1, (uintptr_t)y1on,
                              (uintptr_t)c->rk_sb[0][0], (uintptr_t)c->rk_err[0][0], (uintptr_t)c->rk_sb[1][0],
                              (uintptr_t)c->rk_err[1][0], (uintptr_t)(intptr_t)c->rk_variant};
    std::vector<unsigned char> now((const unsigned char *)key, (const unsigned char *)key + sizeof key);
    if (c->rk_args_dev && now == c->rk_args_key) return LEE_OK;
    if (c->capturing) return set_err(c, LEE_ERR_STATE, "RK4 stage arguments changed inside a captured chunk");
    if (!c->rk_args_dev) {
        int rc = dalloc(c, &c->rk_args_dev, 4 * (sz + szb));
        if (rc) return rc;
    }
    c->rk_args_host.assign(4 * (sz + szb), 0);
    for (int st = 0; st < 4; ++st) {
        RKArgs<T> r;
        RKPoreArgs<T> pr;
        RKBndArgs<T> b;
        rk4_stage_args<T>(c, st, J1, y1on, r, pr, b);
        std::memcpy(c->rk_args_host.data() + st * sz, &r, sz);
        std::memcpy(c->rk_args_host.data() + 4 * sz + st * szb, &b, szb);
    }
    c->rk_args_key.swap(now);
    HIPCHK(c, hipMemcpyAsync(c->rk_args_dev, c->rk_args_host.data(), 4 * (sz + szb), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LEE_OK;
}

// Stage arguments of the coming step, without advancing the gate counter.
template <typename T> int rk_args_peek(lee_ctx *c)
{
    const long long save = c->counter;
    int J1, y1on;
    gate_tick(c, J1, y1on);
    c->counter = save;
    return rk_args_sync<T>(c, J1, y1on);
}

// One RK4 step: launches only (no host synchronisation), receivers into rec_slot.
template <typename T> void enqueue_rk4_step(lee_ctx *c, double *rec_slot)
{
    Model &m = c->m;
    hipStream_t s = c->stream;
    const int nr = (int)m.ri.size();
    const int wind = rk_wind_mode(m);
    int J1, y1on;
    gate_tick(c, J1, y1on);
    if (rk_args_sync<T>(c, J1, y1on)) c->args_fault = true;
    // receivers sample level n before the step (airpore.cpp:573-580)
    launch_sample<T>((const T *)c->air[c->cur * 4 + 3], c->rec_idx, nr, rec_slot, s);
    const RKArgs<T> *rd = (const RKArgs<T> *)c->rk_args_dev;
    const RKBndArgs<T> *bd = (const RKBndArgs<T> *)((const unsigned char *)c->rk_args_dev + 4 * sizeof(RKArgs<T>));
    for (int st = 0; st < 4; ++st) {
        RKArgs<T> r;
        RKPoreArgs<T> pr;
        RKBndArgs<T> b;
        rk4_stage_args<T>(c, st, J1, y1on, r, pr, b);
        prof_mark(c, KS_RK_AIR, true);
        launch_rk_air_v<T>(r, rd + st, wind, m.terrain, c->rk_variant, s);
        prof_mark(c, KS_RK_AIR, false);
        if (m.south_porous) {
            prof_mark(c, KS_RK_PORE, true);
            launch_rk_pore<T>(pr, m.terrain, s);
            prof_mark(c, KS_RK_PORE, false);
        }
        prof_mark(c, KS_RK_BND, true);
        launch_rk_bnd<T>(b, bd + st, s);
        prof_mark(c, KS_RK_BND, false);
    }
    inject_source<T>(c, (T *)c->air[c->cur * 4 + 3]);
    c->steps += 1;
}

// Face-walk arguments of the split FB passes: they take the x-face nodes of the interior rows
// (the face passes then walk only the x planes' edge rings); returns whether they do.
template <typename T> bool fb_face_args(const lee_ctx *c, const AirArgs<T> &a, BCArgs<T> &b)
{
    if (c->fused || !fb_passes_take_xfaces<T>(a)) return false;
    b.xring = 1;
    b.nb_air = boundary_count<T>(c->m.nx, c->m.ny, c->m.nz, true);
    return true;
}

// One FB_p_v step (airpore.cpp:589-614): launches only, receivers into rec_slot.
template <typename T> void enqueue_fb_step(lee_ctx *c, double *rec_slot)
{
    Model &m = c->m;
    hipStream_t s = c->stream;
    int J1, y1on;
    gate_tick(c, J1, y1on);
    AirArgs<T> a = air_args<T>(c, J1, y1on);
    PoreArgs<T> pa{};
    if (m.south_porous) pa = PoreArgs<T>{pore_grid<T>(c), (T)m.p_gama, (T)m.p_invbeta, (T)m.p_eff};
    BCArgs<T> b = bc_args<T>(c, a, rec_slot);
    if (c->fb_form < 0) c->fb_form = (c->fb_mask & 1u) ? 0 : 3;  // not tuned: the path's default
    c->fused = fb_form_fused(c->fb_form);
    const bool pad = c->fb_form == 1, vec_ok = c->fb_form != 6;
    const int per_cu = c->fused ? kFbPerCu[c->fb_form - 2] : 0;
    c->fb_xfaces = fb_face_args<T>(c, a, b);
    if (c->fused) {
        // porous pressure first: the air South face reads its plane Kp-2 through the interface
        if (m.south_porous) {
            prof_mark(c, KS_PORE_P, true);
            launch_pore_p<T>(pa, s);
            prof_mark(c, KS_PORE_P, false);
        }
        prof_mark(c, KS_AIR_FB, true);
        launch_air_fb<T>(a, m.has_wind, s, per_cu);
        prof_mark(c, KS_AIR_FB, false);
        prof_mark(c, KS_BC_P, true);
        launch_bc_p_all<T>(b, s);
        prof_mark(c, KS_BC_P, false);
    } else {
        // ---- pressure phase (airpore.cpp:366-398); receivers sample level n first
        prof_mark(c, KS_AIR_P, true);
        launch_air_p<T>(a, m.has_wind, s, pad, vec_ok);
        prof_mark(c, KS_AIR_P, false);
        if (m.south_porous) {
            prof_mark(c, KS_PORE_P, true);
            launch_pore_p<T>(pa, s);
            prof_mark(c, KS_PORE_P, false);
        }
        prof_mark(c, KS_BC_P, true);
        launch_bc_p_all<T>(b, s);
        prof_mark(c, KS_BC_P, false);
        // ---- velocity phase (airpore.cpp:324-364)
        prof_mark(c, KS_AIR_V, true);
        launch_air_v<T>(a, m.has_wind, s, pad, vec_ok);
        prof_mark(c, KS_AIR_V, false);
    }
    if (m.south_porous) {
        prof_mark(c, KS_PORE_V, true);
        launch_pore_v<T>(pa, s);
        prof_mark(c, KS_PORE_V, false);
    }
    prof_mark(c, KS_BC_V, true);
    launch_bc_v_all<T>(b, s);
    prof_mark(c, KS_BC_V, false);
    if (m.east_tdbc) std::swap(c->psi, c->psi2);
    if (m.ground) std::swap(c->gpsi, c->gpsi2);
    inject_source<T>(c, (T *)c->air[(c->cur ^ 1) * 4 + 3]);
    // ---- level n+1 becomes level n
    c->cur ^= 1;
    c->qcur ^= 1;
    c->steps += 1;
}

template <typename T> void enqueue_step(lee_ctx *c, double *rec_slot)
{
    if (c->m.rk4) enqueue_rk4_step<T>(c, rec_slot);
    else enqueue_fb_step<T>(c, rec_slot);
}

// The y1 gate is a launch argument: a captured chunk must not cross its opening.
bool gate_constant(const lee_ctx *c, int n)
{
    const Model &m = c->m;
    if (!(m.judge[1] == 1 || m.judge[1] == 3)) return true;
    return c->counter + 1 >= m.pml_time || c->counter + n < m.pml_time;
}

void drop_graphs(lee_ctx *c)
{
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

// n steps from a captured graph: found by the buffer roles at its start, else captured now.
// Replaying advances the host-side roles exactly as the capture did.
template <typename T> int graph_steps(lee_ctx *c, int n)
{
    const Model &m = c->m;
    const bool gate = (m.judge[1] == 1 || m.judge[1] == 3);
    const int gate_on = gate && c->counter + 1 >= m.pml_time;
    if (m.rk4) {  // the chunk's stage arguments (constant over it) into device memory first
        int rc = rk_args_peek<T>(c);
        if (rc) return rc;
    }
    for (auto &g : c->graphs)
        if (g.n == n && g.cur == c->cur && g.qcur == c->qcur && g.gate_on == gate_on && g.psi == (void *)c->psi &&
            g.gpsi == (void *)c->gpsi) {
            HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
            if (!m.rk4) {
                if (n & 1) {
                    c->cur ^= 1; c->qcur ^= 1;
                    if (m.east_tdbc) std::swap(c->psi, c->psi2);
                    if (m.ground) std::swap(c->gpsi, c->gpsi2);
                }
            }
            if (gate) c->counter += n;
            c->steps += n;
            return LEE_OK;
        }
    lee_ctx::GraphEntry e{n, c->cur, c->qcur, gate_on, (void *)c->psi, (void *)c->gpsi, nullptr};
    const int nr = (int)m.ri.size();
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    for (int t = 0; t < n; ++t) enqueue_step<T>(c, c->rec_buf + (size_t)t * nr);
    c->capturing = false;
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->stream, &graph));
    if (c->args_fault) {
        (void)hipGraphDestroy(graph);
        c->args_fault = false;
        return set_err(c, LEE_ERR_STATE, "RK4 stage arguments changed inside a captured chunk");
    }
    hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(c, ie);
    if (c->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(c->graphs.front().exec);
        c->graphs.erase(c->graphs.begin());
    }
    c->graphs.push_back(e);
    HIPCHK(c, hipGraphLaunch(e.exec, c->stream));
    return LEE_OK;
}

// Autotuner of the RK4 air stage: times each launch shape on the air and boundary launches of
// stages 0-2 of the coming step (they read level n and the registers stage 0 rewrites, so
// repeating them changes nothing the step itself does not overwrite) and keeps the fastest.
template <typename T> int tune_rk(lee_ctx *c)
{
    const Model &m = c->m;
    if (m.opt.rk_tile > 0) {
        c->rk_variant = std::min(m.opt.rk_tile, kRkVariants) - 1;
        return LEE_OK;
    }
    const long long save = c->counter;
    int J1, y1on;
    gate_tick(c, J1, y1on);
    c->counter = save;
    // the air and boundary launches of stages 0, 1, 2 of the coming step: stage 0 rewrites the
    // accumulators and the stage states that 1 and 2 read (from level n, which none of them
    // writes), so the three can be repeated; stage 3 writes level n and is left out.  The
    // boundary stage is timed with the air stage because its walks depend on the shape (the
    // split shapes take the x-face nodes of the interior rows off it).
    const int wind = rk_wind_mode(m);
    RKArgs<T> r[3];
    RKPoreArgs<T> pr;
    RKBndArgs<T> b[3];
    // two passes over the shapes, each shape's best pass kept: shapes a few percent apart
    // are not ranked by one pass's noise
    float t[kRkVariants];
    for (int v = 0; v < kRkVariants; ++v) t[v] = 1e30f;
    for (int pass = 0; pass < 2; ++pass)
        for (int v = 0; v < kRkVariants; ++v) {
            c->rk_variant = v;  // the stage arguments of this shape (the boundary walks follow it)
            for (int st = 0; st < 3; ++st) rk4_stage_args<T>(c, st, J1, y1on, r[st], pr, b[st]);
            int rc = rk_args_sync<T>(c, J1, y1on);
            if (rc) {
                c->rk_variant = -1;
                return rc;
            }
            const RKArgs<T> *rd = (const RKArgs<T> *)c->rk_args_dev;
            const RKBndArgs<T> *bd =
                (const RKBndArgs<T> *)((const unsigned char *)c->rk_args_dev + 4 * sizeof(RKArgs<T>));
            auto stages = [&]() {
                for (int st = 0; st < 3; ++st) {
                    launch_rk_air_v<T>(r[st], rd + st, wind, m.terrain, v, c->stream);
                    launch_rk_bnd<T>(b[st], bd + st, c->stream);
                }
            };
            stages();  // warm-up
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            for (int rep = 0; rep < 2; ++rep) stages();
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev1));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            t[v] = std::min(t[v], ms);
        }
    HIPCHK(c, hipGetLastError());
    int pick = 0;
    for (int v = 1; v < kRkVariants; ++v)
        if (t[v] < t[pick]) pick = v;
    c->rk_variant = pick;
    return LEE_OK;
}

// Autotuner of the FB_p_v air passes: times each launch form the path allows on the air and
// face passes of the coming step (they write only level n+1, its auxiliaries and the
// receiver buffer, all of which the step itself rewrites) and keeps the fastest.  The porous
// passes are the same in every form and are left out.
template <typename T> int tune_fb(lee_ctx *c)
{
    const Model &m = c->m;
    const long long save = c->counter;
    int J1, y1on;
    gate_tick(c, J1, y1on);
    c->counter = save;
    const AirArgs<T> a = air_args<T>(c, J1, y1on);
    float t[kFbForms];
    for (int f = 0; f < kFbForms; ++f) t[f] = 1e30f;
    for (int pass = 0; pass < 2; ++pass)
        for (int f = 0; f < kFbForms; ++f) {
            if (!(c->fb_mask >> f & 1u)) continue;
            c->fused = fb_form_fused(f);
            BCArgs<T> b = bc_args<T>(c, a, c->rec_buf);
            (void)fb_face_args<T>(c, a, b);
            auto passes = [&]() {
                if (c->fused) {
                    launch_air_fb<T>(a, m.has_wind, c->stream, kFbPerCu[f - 2]);
                    launch_bc_p_all<T>(b, c->stream);
                } else {
                    launch_air_p<T>(a, m.has_wind, c->stream, f == 1, f != 6);
                    launch_bc_p_all<T>(b, c->stream);
                    launch_air_v<T>(a, m.has_wind, c->stream, f == 1, f != 6);
                }
                launch_bc_v_all<T>(b, c->stream);
            };
            passes();  // warm-up
            HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            for (int rep = 0; rep < 3; ++rep) passes();
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev1));
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            t[f] = std::min(t[f], ms);
        }
    HIPCHK(c, hipGetLastError());
    int pick = -1;
    for (int f = 0; f < kFbForms; ++f)
        if ((c->fb_mask >> f & 1u) && (pick < 0 || t[f] < t[pick])) pick = f;
    c->fb_form = pick;
    c->fused = fb_form_fused(pick);
    return LEE_OK;
}

template <typename T> int run_steps(lee_ctx *c, int nsteps, double *rec_out)
{
    Model &m = c->m;
    hipStream_t s = c->stream;
    if (m.rk4 && c->rk_variant < 0) {
        int rc = tune_rk<T>(c);
        if (rc) return rc;
    }
    if (!m.rk4 && c->fb_form < 0) {
        int rc = tune_fb<T>(c);
        if (rc) return rc;
    }
    const int nr = (int)m.ri.size();
    HIPCHK(c, hipEventRecord(c->ev0, s));
    int done = 0;
    while (done < nsteps) {
        int n = std::min(nsteps - done, kRecChunk);
        const long long before = c->steps;
        if (m.opt.use_graph && !c->profile && c->src_ps.empty()) {
            n = std::min(n, kGraphSteps);
            if (gate_constant(c, n)) {
                int rc = graph_steps<T>(c, n);
                if (rc) return rc;
            } else {
                for (int t = 0; t < n; ++t) enqueue_step<T>(c, c->rec_buf + (size_t)t * nr);
            }
        } else {
            for (int t = 0; t < n; ++t) enqueue_step<T>(c, c->rec_buf + (size_t)t * nr);
        }
        if (c->args_fault) {
            c->args_fault = false;
            return set_err(c, LEE_ERR_DEVICE, "RK4 stage arguments could not be uploaded: " + c->err);
        }
        HIPCHK(c, hipGetLastError());
        if (rec_out && nr) {
            HIPCHK(c, hipMemcpyAsync(rec_out + (size_t)done * nr, c->rec_buf, (size_t)n * nr * 8, hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipStreamSynchronize(s));
        }
        done += n;
        const int he = m.opt.health_every;
        if (he > 0 && c->steps / he != before / he) {
            HIPCHK(c, hipMemsetAsync(c->flag, 0, sizeof(int), s));
            launch_finite_check<T>((const T *)c->air[c->cur * 4 + 3], air_n(c), c->flag, s);
            int h = 0;
            HIPCHK(c, hipMemcpyAsync(&h, c->flag, sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipStreamSynchronize(s));
            if (h) return set_err(c, LEE_ERR_NONFINITE, "non-finite pressure after step " + std::to_string(c->steps));
        }
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, s));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    HIPCHK(c, hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    prof_resolve(c);
    return LEE_OK;
}

int create_impl(const lee_deck *deck, const lee_options *opt, lee_ctx **out)
{
    std::unique_ptr<lee_ctx> c(new lee_ctx());
    int rc0;
    lee_options o;
    lee_options_default(&o);
    if (opt) o = *opt;
    if (o.precision != 64 && o.precision != 32) return set_err(nullptr, LEE_ERR_ARG, "precision must be 32 or 64");
    int code = LEE_OK;
    std::string e = c->m.build(*deck, o, code);
    if (!e.empty()) return set_err(nullptr, code, e);
    lee_ctx *C = c.get();
    C->prec = o.precision;
    C->esz = o.precision == 64 ? 8 : 4;
    {
        const Model &mm = C->m;
        const bool can = fused_fits<double>(mm.nx) && !(mm.south_porous && !mm.iface_copy) && !mm.terrain && !mm.rk4 &&
                         !mm.obstacle && !mm.ground;
        if (o.path == LEE_PATH_FUSED && !can)
            return set_err(nullptr, LEE_ERR_UNSUPPORTED,
                           "the fused path runs the FB_p_v step on a separable grid and, with a porous South face, "
                           "needs the interface copy");
        if (o.path < LEE_PATH_AUTO || o.path > LEE_PATH_FUSED) return set_err(nullptr, LEE_ERR_ARG, "unknown path");
        if (o.rk_tile < 0 || o.rk_tile > kRkVariants) return set_err(nullptr, LEE_ERR_ARG, "rk_tile must be 0..11");
        // the FB launch forms the path allows: SPLIT the split passes (plain, padded, or with
        // the one-point interior), FUSED the fused pass at each pencil depth, AUTO both when
        // the fused pass can run
        C->fb_mask = o.path == LEE_PATH_FUSED ? kFbFusedForms
                     : o.path == LEE_PATH_SPLIT || !can ? kFbSplitForms : kFbSplitForms | kFbFusedForms;
        if (o.fb_form < 0 || o.fb_form > kFbForms) return set_err(nullptr, LEE_ERR_ARG, "fb_form must be 0..7");
        if (o.fb_form > 0) {
            if (!(C->fb_mask >> (o.fb_form - 1) & 1u))
                return set_err(nullptr, LEE_ERR_UNSUPPORTED, "fb_form is not one of the forms this path and deck allow");
            C->fb_form = o.fb_form - 1;
        }
        C->fused = fb_form_fused(C->fb_form) || o.path == LEE_PATH_FUSED;
        // the boundary kernels decode node indices in 32 bits
        const long long nb = boundary_count<double>(mm.nx, mm.ny, mm.nz) +
                             (mm.south_porous ? boundary_count<double>(mm.pnx, mm.pny, mm.pnz) : 0) + (long long)mm.ri.size();
        if (nb >= (1LL << 31)) return set_err(nullptr, LEE_ERR_ARG, "grid too large: more than 2^31 boundary nodes");
    }
    HIPCHK(C, hipSetDevice(o.device));
    C->device = o.device;
    if ((rc0 = scratch_setup(C, o.device))) return rc0;
    HIPCHK(C, hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
    HIPCHK(C, hipEventCreate(&C->ev0));
    HIPCHK(C, hipEventCreate(&C->ev1));
    const Model &m = C->m;
    size_t na = air_n(C), np = pore_n(C), es = C->esz;
    int rc;
    for (int f = 0; f < 8; ++f)
        if ((rc = dalloc(C, &C->air[f], na * es))) return rc;
    if (m.south_porous)
        for (int f = 0; f < 8; ++f)
            if ((rc = dalloc(C, &C->pore[f], np * es))) return rc;
    if (m.has_wind)
        if ((rc = dalloc(C, &C->U, na * es)) || (rc = dalloc(C, &C->V, na * es)) || (rc = dalloc(C, &C->W, na * es))) return rc;
    if ((rc = dalloc(C, &C->mx, m.nx * es)) || (rc = dalloc(C, &C->my, m.ny * es)) || (rc = dalloc(C, &C->mz, m.nz * es))) return rc;
    if (m.south_porous)
        if ((rc = dalloc(C, &C->pmx, m.pnx * es)) || (rc = dalloc(C, &C->pmy, m.pny * es)) || (rc = dalloc(C, &C->pmz, m.pnz * es)))
            return rc;
    int len[6] = {m.nx, m.nx, m.ny, m.ny, m.nz, m.nz};
    size_t qmax = 0;
    for (int s = 0; s < 6; ++s) {
        for (int k = 0; k < 3; ++k)
            if ((rc = dalloc(C, &C->prof[s][k], (size_t)len[s] * es))) return rc;
        size_t cnt = m.slab[s].count(m.nx, m.ny, m.nz);
        qmax = std::max(qmax, cnt);
        for (int h = 0; h < 2; ++h)
            for (int f = 0; f < 4; ++f)
                if ((rc = dalloc(C, &C->Q[h][s][f], cnt * es))) return rc;
    }
    if ((rc = dalloc(C, (void **)&C->psi, (size_t)m.nx * m.nz * 6 * 8))) return rc;
    if ((rc = dalloc(C, (void **)&C->psi2, (size_t)m.nx * m.nz * 6 * 8))) return rc;
    if (m.ground) {
        const size_t ng = (size_t)m.nx * m.ny * m.gK * 8;
        if ((rc = dalloc(C, (void **)&C->gpsi, ng)) || (rc = dalloc(C, (void **)&C->gpsi2, ng))) return rc;
        if ((rc = dalloc(C, (void **)&C->gmx, (size_t)m.nx * 8)) || (rc = dalloc(C, (void **)&C->gmy, (size_t)m.ny * 8)))
            return rc;
        if (m.terrain && (rc = dalloc(C, (void **)&C->gter, (size_t)2 * m.nx * m.ny * 8))) return rc;
    }
    int nr = (int)m.ri.size();
    if ((rc = dalloc(C, (void **)&C->rec_idx, (size_t)std::max(nr, 1) * sizeof(long long)))) return rc;
    if ((rc = dalloc(C, (void **)&C->rec_buf, (size_t)kRecChunk * std::max(nr, 1) * 8))) return rc;
    if ((rc = dalloc(C, (void **)&C->flag, sizeof(int)))) return rc;
    if (m.rk4) {
        for (int d = 0; d < (m.south_porous ? 2 : 1); ++d) {
            size_t n = d ? np : na;
            // the air accumulators of the four fields side by side (acc4_load / acc4_store), the
            // porous ones a field each
            if (d == 0 && (rc = dalloc(C, &C->rk_acc[0][0], 4 * n * es))) return rc;
            for (int f = 0; f < 4; ++f) {
                if ((rc = dalloc(C, &C->rk_sb[d][f], n * es)) || (d == 1 && (rc = dalloc(C, &C->rk_acc[d][f], n * es))))
                    return rc;
                // FP32: the carried state error; the low part of the accumulator only for the
                // porous velocities (integrating-factor update)
                if (C->prec == 32) {
                    if ((rc = dalloc(C, &C->rk_err[d][f], n * es))) return rc;
                    if (d == 1 && f < 3 && (rc = dalloc(C, &C->rk_lo[d][f], n * es))) return rc;
                }
            }
        }
        for (int s = 0; s < 6; ++s) {
            size_t cnt = m.slab[s].count(m.nx, m.ny, m.nz);
            for (int f = 0; f < 4; ++f)
                if ((rc = dalloc(C, &C->rkq_sb[s][f], cnt * es))) return rc;
            // the accumulators of the four fields side by side (acc4_load / acc4_store)
            if ((rc = dalloc(C, &C->rkq_acc[s][0], 4 * cnt * es))) return rc;
        }
    }
    if (m.terrain) {
        for (int d = 0; d < (m.south_porous ? 2 : 1); ++d) {
            size_t n2 = d ? (size_t)m.pnx * m.pny : (size_t)m.nx * m.ny;
            size_t nk = d ? (size_t)m.pnz : (size_t)m.nz;
            for (int q = 0; q < 3; ++q)
                if ((rc = dalloc(C, &C->ter[d][q], n2 * es))) return rc;
            if ((rc = dalloc(C, &C->ter[d][3], nk * es))) return rc;
        }
    }
    if (m.has_wind && m.wind_v_only) {
        if ((rc = dalloc(C, (void **)&C->wind_z, (size_t)2 * m.nz * 8))) return rc;
        if (m.terrain && (rc = dalloc(C, (void **)&C->wind_h, (size_t)m.nx * m.ny * 8))) return rc;
        std::vector<double> zt(2 * m.nz, 0.0);
        if (m.terrain) {
            FrameTables f0;
            m.frame_tables(0, f0, false);
            for (int k = 0; k < m.nz; ++k) zt[m.nz + k] = f0.ter.tk[k];
        }
        for (int k = 0; k < m.nz; ++k) zt[k] = m.az[k];
        if ((rc = upload_f64(C, C->wind_z, zt.data(), zt.size()))) return rc;
    }
    if (m.obstacle) {
        if ((rc = dalloc(C, (void **)&C->omask, m.omask.size()))) return rc;
        HIPCHK(C, hipMemcpyAsync(C->omask, m.omask.data(), m.omask.size(), hipMemcpyHostToDevice, C->stream));
        HIPCHK(C, hipStreamSynchronize(C->stream));
    }
    C->shift_tmp_bytes = std::max<size_t>(qmax * es, 8);
    if ((rc = dalloc(C, &C->shift_tmp, C->shift_tmp_bytes))) return rc;
    // initial condition (calculation_2D.cpp:300-314): velocities zero, Gaussian pressure
    std::vector<double> p0;
    m.initial_condition(p0, false);
    if ((rc = upload(C, C->air[3], p0.data(), p0.size()))) return rc;
    if (m.south_porous) {
        m.initial_condition(p0, true);
        if ((rc = upload(C, C->pore[3], p0.data(), p0.size()))) return rc;
    }
    HIPCHK(C, hipStreamSynchronize(C->stream));
    *out = c.release();
    return LEE_OK;
}

}  // namespace

extern "C" {

const char *lee_global_error(void) { return g_global_err.c_str(); }

const char *lee_last_error(const lee_ctx *ctx) { return ctx ? ctx->err.c_str() : g_global_err.c_str(); }

int lee_create(const lee_deck *deck, const lee_options *opt, lee_ctx **out)
{
    if (!deck || !out) return set_err(nullptr, LEE_ERR_ARG, "NULL argument");
    *out = nullptr;
    try {
        return create_impl(deck, opt, out);
    } catch (const std::exception &e) {
        return set_err(nullptr, LEE_ERR_ARG, std::string("lee_create: ") + e.what());
    }
}

int lee_destroy(lee_ctx *ctx)
{
    delete ctx;
    return LEE_OK;
}

int lee_get_info(const lee_ctx *c, lee_info *info)
{
    if (!c || !info) return LEE_ERR_ARG;
    const Model &m = c->m;
    info->nx = m.nx; info->ny = m.ny; info->nz = m.nz;
    info->pnx = m.pnx; info->pny = m.pny; info->pnz = m.pnz;
    info->ijkmax = m.ijkmax;
    info->n_receivers = (int)m.ri.size();
    info->precision = c->prec;
    info->mf_count = c->mf_count;
    info->steps_done = c->steps;
    info->pml_y1_counter = c->counter;
    info->path = c->fused ? LEE_PATH_FUSED : LEE_PATH_SPLIT;
    info->rk_tile = c->rk_variant + 1;
    info->fb_form = c->m.rk4 ? 0 : c->fb_form + 1;
    return LEE_OK;
}

int lee_receiver_indices(const lee_ctx *c, int *ijk, int cap)
{
    if (!c || !ijk) return LEE_ERR_ARG;
    int n = (int)c->m.ri.size();
    for (int r = 0; r < n && 3 * r + 2 < cap; ++r) {
        ijk[3 * r] = c->m.ri[r]; ijk[3 * r + 1] = c->m.rj[r]; ijk[3 * r + 2] = c->m.rk[r];
    }
    return n;
}

int lee_frame_setup(lee_ctx *c, int N)
{
    if (!c || N < 0) return set_err(c, LEE_ERR_ARG, "bad frame number");
    if (N >= c->m.frame_count()) return set_err(c, LEE_ERR_ARG, "frame number beyond the last frame");
    drop_graphs(c);  // a frame shift changes buffer roles the graphs were captured with
    int rc = c->prec == 64 ? frame_shift<double>(c, N) : frame_shift<float>(c, N);
    if (rc) return rc;
    if (N == 0)  // air::SetQMove(0) zeroes every auxiliary (air.cpp:2453-2519)
        for (int h = 0; h < 2; ++h)
            for (int s = 0; s < 6; ++s)
                for (int f = 0; f < 4; ++f) {
                    size_t cnt = c->m.slab[s].count(c->m.nx, c->m.ny, c->m.nz);
                    if (cnt) HIPCHK(c, hipMemsetAsync(c->Q[h][s][f], 0, cnt * c->esz, c->stream));
                }
    rc = upload_frame(c, N);
    if (rc) return rc;
    c->mf_count = N;
    return LEE_OK;
}

int lee_step(lee_ctx *c, int nsteps, double *rec_out)
{
    if (!c || nsteps < 0) return set_err(c, LEE_ERR_ARG, "bad arguments");
    if (c->mf_count < 0) return set_err(c, LEE_ERR_STATE, "lee_frame_setup must run before lee_step");
    if (nsteps == 0) return LEE_OK;
    return c->prec == 64 ? run_steps<double>(c, nsteps, rec_out) : run_steps<float>(c, nsteps, rec_out);
}

int lee_set_source_series(lee_ctx *c, const double *ps, long long n, int period)
{
    if (!c || n < 0 || (n > 0 && (!ps || period < 1))) return set_err(c, LEE_ERR_ARG, "bad source series");
    if (n > 0 && c->m.terrain) return set_err(c, LEE_ERR_UNSUPPORTED, "the time-series source is defined on the separable grid");
    for (long long q = 0; q < n; ++q)
        if (!std::isfinite(ps[q])) return set_err(c, LEE_ERR_ARG, "non-finite source amplitude");
    const Model &m = c->m;
    if (n > 0 && !c->src_xyz) {
        int rc = dalloc(c, (void **)&c->src_xyz, (size_t)(m.nx + m.ny + m.nz) * 8);
        if (rc) return rc;
        if ((rc = upload_f64(c, c->src_xyz, m.ax.data(), m.nx))) return rc;
        if ((rc = upload_f64(c, c->src_xyz + m.nx + m.ny, m.az.data(), m.nz))) return rc;
        FrameTables t;
        m.frame_tables(c->mf_count < 0 ? 0 : c->mf_count, t, false);
        if ((rc = upload_f64(c, c->src_xyz + m.nx, t.Y.data(), t.Y.size()))) return rc;
    }
    c->src_ps.assign(ps, ps + n);
    c->src_period = n > 0 ? period : 0;
    drop_graphs(c);
    return LEE_OK;
}

long long lee_read_source_series(const char *dir, double *ps, long long cap)
{
    if (!dir || cap < 0 || (cap > 0 && !ps)) { set_global_error("bad arguments"); return LEE_ERR_ARG; }
    std::string path = std::string(dir) + "/source_input.dat";
    std::ifstream in(path);
    if (!in) { set_global_error("cannot open " + path); return LEE_ERR_IO; }
    long long n = 0;
    double v;
    while (n < cap && (in >> v)) ps[n++] = v;
    return n;
}

int lee_frame_steps(const lee_ctx *c, int mf, int temp_time) { return c ? c->m.frame_steps(mf, temp_time) : 0; }
int lee_frame_count(const lee_ctx *c) { return c ? c->m.frame_count() : 0; }

int lee_run(lee_ctx *c, double *rec_out, long long cap)
{
    if (!c) return set_err(c, LEE_ERR_ARG, "NULL context");
    const Model &m = c->m;
    int nr = (int)m.ri.size();
    if (rec_out && cap < (long long)m.d.time_domain * nr) return set_err(c, LEE_ERR_ARG, "receiver buffer too small");
    int mf = 0, temp_time = 0, last = m.last_frame();
    float total = 0.f;
    do {
        int lim = m.frame_steps(mf, temp_time);
        int rc = lee_frame_setup(c, mf);
        if (rc) return rc;
        if (lim > 0) {
            rc = lee_step(c, lim, rec_out ? rec_out + (size_t)temp_time * nr : nullptr);
            if (rc) return rc;
            total += c->last_ms;
        }
        mf += 1;
        temp_time += lim > 0 ? lim : 0;
    } while (mf <= last && mf < m.frame_count() && temp_time < m.d.time_domain && m.d.mf_judge == 1);
    c->last_ms = total;
    return LEE_OK;
}

int lee_get_state(lee_ctx *c, int domain, double *u, double *v, double *w, double *p)
{
    if (!c || (domain != 0 && domain != 1)) return set_err(c, LEE_ERR_ARG, "bad arguments");
    if (domain == 1 && !c->m.south_porous) return set_err(c, LEE_ERR_ARG, "no porous domain");
    void **f = domain == 0 ? c->air : c->pore;
    size_t n = domain == 0 ? air_n(c) : pore_n(c);
    int b = c->cur * 4, rc;
    double *dst[4] = {u, v, w, p};
    for (int q = 0; q < 4; ++q)
        if ((rc = download(c, dst[q], f[b + q], n))) return rc;
    return LEE_OK;
}

int lee_set_state(lee_ctx *c, int domain, const double *u, const double *v, const double *w, const double *p)
{
    if (!c || (domain != 0 && domain != 1)) return set_err(c, LEE_ERR_ARG, "bad arguments");
    if (domain == 1 && !c->m.south_porous) return set_err(c, LEE_ERR_ARG, "no porous domain");
    void **f = domain == 0 ? c->air : c->pore;
    size_t n = domain == 0 ? air_n(c) : pore_n(c);
    int b = c->cur * 4, rc;
    const double *src[4] = {u, v, w, p};
    for (int q = 0; q < 4; ++q)
        if (src[q] && (rc = upload(c, f[b + q], src[q], n))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LEE_OK;
}

}  // extern "C"

// ---- checkpoint / restart in the reference's text format -------------------------------
namespace {

// Device array of n elements of the context precision -> FP64 host vector.
int fetch(lee_ctx *c, const void *src, size_t n, std::vector<double> &out)
{
    out.assign(n, 0.0);
    if (!src || !n) return LEE_OK;
    return download(c, out.data(), src, n);
}

// The stream format of calculation_2D::save_restart_cal / air::save_restart_air
// (calculation_2D.cpp:497-520, air.cpp:1305-1390): scientific, explicit sign, 10 digits,
// width 18 for the first item only, one space after every value, a newline per section.
void restart_stream(std::ofstream &o)
{
    o.setf(std::ios::scientific, std::ios::floatfield);
    o.setf(std::ios::showpos);
    o.precision(10);
    o.width(18);
}

}  // namespace

extern "C" {

// Reference layout: f[i][j][k] loops (i outer, k inner) for u v w p; then U V W of the mean
// flow; then the PML slabs x1, x2 (i, j, k loops), y1, y2 (j, k, i loops), z1, z2 (k, i, j
// loops), each Q_u Q_v Q_w Q_p per node; then the porous u v w p (airpore.cpp:783-796).
int lee_save_restart(lee_ctx *c, const char *path, int index)
{
    if (!c || !path) return set_err(c, LEE_ERR_ARG, "bad arguments");
    const Model &m = c->m;
    int rc;
    std::vector<double> f[4], U, V, W;
    const int b = c->cur * 4;
    for (int q = 0; q < 4; ++q)
        if ((rc = fetch(c, c->air[b + q], air_n(c), f[q]))) return rc;
    if ((rc = fetch(c, m.has_wind ? c->U : nullptr, m.has_wind ? air_n(c) : 0, U)) ||
        (rc = fetch(c, m.has_wind ? c->V : nullptr, m.has_wind ? air_n(c) : 0, V)) ||
        (rc = fetch(c, m.has_wind ? c->W : nullptr, m.has_wind ? air_n(c) : 0, W)))
        return rc;
    {
        // header: mpi_size mpi_rank index (airpore.cpp:686-688)
        std::ofstream h(path, std::ios::out | std::ios::binary | std::ios::trunc);
        if (!h) return set_err(c, LEE_ERR_IO, std::string("cannot write ") + path);
        h << 1 << " " << 0 << " " << index << std::endl;
    }
    std::ofstream o(path, std::ios::app | std::ios::binary);
    if (!o) return set_err(c, LEE_ERR_IO, std::string("cannot write ") + path);
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    auto id = [&](int i, int j, int k) { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k); };
    restart_stream(o);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                for (int q = 0; q < 4; ++q) o << f[q][id(i, j, k)] << " ";
    o << std::endl;
    restart_stream(o);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const size_t x = id(i, j, k);
                o << (U.empty() ? 0.0 : U[x]) << " " << (V.empty() ? 0.0 : V[x]) << " " << (W.empty() ? 0.0 : W[x]) << " ";
            }
    std::vector<double> Q[4];
    for (int s = 0; s < 6; ++s) {
        const SlabGeom &g = m.slab[s];
        const size_t cnt = g.count(nx, ny, nz);
        for (int q = 0; q < 4; ++q)
            if ((rc = fetch(c, c->Q[c->qcur][s][q], cnt, Q[q]))) return rc;
        const int n = g.n;
        if (!n) continue;
        if (g.axis == 0) {
            for (int ii = 0; ii < n; ++ii)
                for (int j = 0; j < ny; ++j)
                    for (int k = 0; k < nz; ++k) {
                        const size_t x = (size_t)ii + (size_t)n * ((size_t)j + (size_t)ny * k);
                        for (int q = 0; q < 4; ++q) o << Q[q][x] << " ";
                    }
        } else if (g.axis == 1) {
            for (int jj = 0; jj < n; ++jj)
                for (int k = 0; k < nz; ++k)
                    for (int i = 0; i < nx; ++i) {
                        const size_t x = (size_t)i + (size_t)nx * ((size_t)jj + (size_t)n * k);
                        for (int q = 0; q < 4; ++q) o << Q[q][x] << " ";
                    }
        } else {
            for (int kk = 0; kk < n; ++kk)
                for (int i = 0; i < nx; ++i)
                    for (int j = 0; j < ny; ++j) {
                        const size_t x = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * kk);
                        for (int q = 0; q < 4; ++q) o << Q[q][x] << " ";
                    }
        }
    }
    o << std::endl;
    if (m.south_porous) {
        std::vector<double> pf[4];
        for (int q = 0; q < 4; ++q)
            if ((rc = fetch(c, c->pore[b + q], pore_n(c), pf[q]))) return rc;
        restart_stream(o);
        const int px = m.pnx, py = m.pny, pz = m.pnz;
        for (int i = 0; i < px; ++i)
            for (int j = 0; j < py; ++j)
                for (int k = 0; k < pz; ++k)
                    for (int q = 0; q < 4; ++q) o << pf[q][(size_t)i + (size_t)px * ((size_t)j + (size_t)py * k)] << " ";
        o << std::endl;
    }
    // extension block (not in the reference, whose reader stops before it): frame number,
    // y1 gate counter, steps done and the East TDBC histories
    o.unsetf(std::ios::showpos);
    o << "LEE-EXT " << c->mf_count << " " << c->counter << " " << c->steps << std::endl;
    const size_t np = (size_t)nx * nz * 6;
    std::vector<double> psi(np);
    HIPCHK(c, hipMemcpyAsync(psi.data(), c->psi, np * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    restart_stream(o);
    for (size_t q = 0; q < np; ++q) o << psi[q] << " ";
    o << std::endl;
    if (m.ground) {  // ground impedance histories
        const size_t ng = (size_t)nx * m.ny * m.gK;
        std::vector<double> d(ng), g(ng);
        HIPCHK(c, hipMemcpyAsync(d.data(), c->gpsi, ng * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        ground_psi_layout(d.data(), g.data(), (size_t)nx * m.ny, m.gK, true);
        for (size_t q = 0; q < ng; ++q) o << g[q] << " ";
        o << std::endl;
    }
    if (!o) return set_err(c, LEE_ERR_IO, std::string("write failed: ") + path);
    return LEE_OK;
}

// Reads what lee_save_restart writes (and the reference's files, which end before the
// extension block: then frame, counter and histories keep their values).  The moving
// frame of the context must already be the file's (lee_frame_setup).
int lee_load_restart(lee_ctx *c, const char *path, int *index)
{
    if (!c || !path) return set_err(c, LEE_ERR_ARG, "bad arguments");
    const Model &m = c->m;
    std::ifstream in(path, std::ios::in | std::ios::binary);
    if (!in) return set_err(c, LEE_ERR_IO, std::string("cannot open ") + path);
    int size = 0, rank = 0, idx = 0;
    if (!(in >> size >> rank >> idx)) return set_err(c, LEE_ERR_IO, "malformed restart header");
    if (size != 1) return set_err(c, LEE_ERR_ARG, "restart written by " + std::to_string(size) + " ranks; one domain expected");
    const int nx = m.nx, ny = m.ny, nz = m.nz;
    auto id = [&](int i, int j, int k) { return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k); };
    std::vector<double> f[4];
    for (int q = 0; q < 4; ++q) f[q].assign(air_n(c), 0.0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k)
                for (int q = 0; q < 4; ++q)
                    if (!(in >> f[q][id(i, j, k)])) return set_err(c, LEE_ERR_IO, "restart file ends inside the fields");
    double dummy;
    for (size_t x = 0; x < 3 * air_n(c); ++x)  // the mean flow is rebuilt from the deck
        if (!(in >> dummy)) return set_err(c, LEE_ERR_IO, "restart file ends inside the mean flow");
    int rc;
    for (int s = 0; s < 6; ++s) {
        const SlabGeom &g = m.slab[s];
        const int n = g.n;
        if (!n) continue;
        const size_t cnt = g.count(nx, ny, nz);
        std::vector<double> Q[4];
        for (int q = 0; q < 4; ++q) Q[q].assign(cnt, 0.0);
        auto rd = [&](size_t x) {
            for (int q = 0; q < 4; ++q)
                if (!(in >> Q[q][x])) return false;
            return true;
        };
        bool ok = true;
        if (g.axis == 0) {
            for (int ii = 0; ii < n && ok; ++ii)
                for (int j = 0; j < ny && ok; ++j)
                    for (int k = 0; k < nz && ok; ++k) ok = rd((size_t)ii + (size_t)n * ((size_t)j + (size_t)ny * k));
        } else if (g.axis == 1) {
            for (int jj = 0; jj < n && ok; ++jj)
                for (int k = 0; k < nz && ok; ++k)
                    for (int i = 0; i < nx && ok; ++i) ok = rd((size_t)i + (size_t)nx * ((size_t)jj + (size_t)n * k));
        } else {
            for (int kk = 0; kk < n && ok; ++kk)
                for (int i = 0; i < nx && ok; ++i)
                    for (int j = 0; j < ny && ok; ++j) ok = rd((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * kk));
        }
        if (!ok) return set_err(c, LEE_ERR_IO, "restart file ends inside the PML auxiliaries");
        for (int q = 0; q < 4; ++q)
            if ((rc = upload(c, c->Q[c->qcur][s][q], Q[q].data(), cnt))) return rc;
    }
    const int b = c->cur * 4;
    for (int q = 0; q < 4; ++q)
        if ((rc = upload(c, c->air[b + q], f[q].data(), air_n(c)))) return rc;
    if (m.south_porous) {
        const int px = m.pnx, py = m.pny, pz = m.pnz;
        std::vector<double> pf[4];
        for (int q = 0; q < 4; ++q) pf[q].assign(pore_n(c), 0.0);
        for (int i = 0; i < px; ++i)
            for (int j = 0; j < py; ++j)
                for (int k = 0; k < pz; ++k)
                    for (int q = 0; q < 4; ++q)
                        if (!(in >> pf[q][(size_t)i + (size_t)px * ((size_t)j + (size_t)py * k)]))
                            return set_err(c, LEE_ERR_IO, "restart file ends inside the porous fields");
        for (int q = 0; q < 4; ++q)
            if ((rc = upload(c, c->pore[b + q], pf[q].data(), pore_n(c)))) return rc;
    }
    std::string tag;
    if (in >> tag && tag == "LEE-EXT") {
        int mf = 0;
        long long counter = 0, steps = 0;
        if (!(in >> mf >> counter >> steps)) return set_err(c, LEE_ERR_IO, "malformed extension block");
        const size_t np = (size_t)nx * nz * 6;
        std::vector<double> psi(np);
        for (size_t q = 0; q < np; ++q)
            if (!(in >> psi[q])) return set_err(c, LEE_ERR_IO, "restart file ends inside the TDBC histories");
        HIPCHK(c, hipMemcpyAsync(c->psi, psi.data(), np * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (m.ground) {
            const size_t ng = (size_t)nx * m.ny * m.gK;
            std::vector<double> g(ng), d(ng);
            for (size_t q = 0; q < ng; ++q)
                if (!(in >> g[q])) return set_err(c, LEE_ERR_IO, "restart file ends inside the ground histories");
            ground_psi_layout(g.data(), d.data(), (size_t)nx * m.ny, m.gK, false);
            HIPCHK(c, hipMemcpyAsync(c->gpsi, d.data(), ng * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        c->counter = counter;
        c->steps = steps;
        if (mf != c->mf_count) return set_err(c, LEE_ERR_STATE, "restart written in frame " + std::to_string(mf) +
                                                                  ", context is in frame " + std::to_string(c->mf_count));
    }
    // the FP32 RK4 state error restarts at zero
    for (int d = 0; d < 2; ++d)
        for (int q = 0; q < 4; ++q)
            if (c->rk_err[d][q]) HIPCHK(c, hipMemsetAsync(c->rk_err[d][q], 0, (d ? pore_n(c) : air_n(c)) * c->esz, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (index) *index = idx;
    return LEE_OK;
}

}  // extern "C"

// ---- contour output: airpore::get_data_contour (airpore.cpp:1018-1050) ------------------
namespace {

int out_count(int n, int stride) { return (n - 1) % stride == 0 ? (n - 1) / stride + 1 : (n - 1) / stride + 2; }

// calculation_2D.cpp:207-213 (out_n = 4) and the subsample of mpi_send_data
// (calculation_2D.cpp:1030-1105); getOstream (calculation_2D.cpp:1165-1182) prints it.
// The zone's subsample goes to `stage` when given (lee_write_contour_noalloc: the GPU's
// staging array, copied back in fixed-size chunks through a stack buffer) or to a device
// buffer allocated for the call (lee_write_contour).
template <typename T>
int contour_zone(lee_ctx *c, std::ofstream &o, const T *p, int nx, int ny, int nz, const std::vector<double> &X,
                 const std::vector<double> &Y, const std::vector<double> &Z, double yshift, double *stage)
{
    const int st = 4;
    const int ox = out_count(nx, st), oy = out_count(ny, st), oz = out_count(nz, st);
    const size_t n = (size_t)ox * oy * oz;
    auto pick = [&](const std::vector<double> &v, int o_, int nn, int q) { return q < o_ - 1 ? v[(size_t)q * st] : v[nn - 1]; };
    o << "I=" << ox << ",J=" << oy << ",K=" << oz << ",F=POINT" << std::endl;
    if (stage) {
        launch_subsample<T>(p, nx, ny, nz, st, ox, oy, oz, stage, c->stream);
        HIPCHK(c, hipGetLastError());
        constexpr size_t kChunk = 1024;
        double h[kChunk];
        size_t r = 0;
        for (int k = 0; k < oz; ++k)
            for (int j = 0; j < oy; ++j)
                for (int i = 0; i < ox; ++i, ++r) {
                    if (r % kChunk == 0) {
                        const size_t cnt = std::min(kChunk, n - r);
                        HIPCHK(c, hipMemcpyAsync(h, stage + r, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                        HIPCHK(c, hipStreamSynchronize(c->stream));
                    }
                    o << pick(X, ox, nx, i) << " " << pick(Y, oy, ny, j) + yshift << " " << pick(Z, oz, nz, k) << " "
                      << h[r % kChunk] << " " << std::endl;
                }
        return LEE_OK;
    }
    double *d = nullptr;
    HIPCHK(c, hipMalloc(&d, n * sizeof(double)));
    launch_subsample<T>(p, nx, ny, nz, st, ox, oy, oz, d, c->stream);
    std::vector<double> h(n);
    hipError_t e1 = hipMemcpyAsync(h.data(), d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    HIPCHK(c, e1);
    HIPCHK(c, e2);
    size_t r = 0;
    for (int k = 0; k < oz; ++k)
        for (int j = 0; j < oy; ++j)
            for (int i = 0; i < ox; ++i, ++r)
                o << pick(X, ox, nx, i) << " " << pick(Y, oy, ny, j) + yshift << " " << pick(Z, oz, nz, k) << " " << h[r]
                  << " " << std::endl;
    return LEE_OK;
}

size_t contour_values(int nx, int ny, int nz)
{
    return (size_t)out_count(nx, 4) * out_count(ny, 4) * out_count(nz, 4);
}

// Writes dir/ptNNN.dat, NNN = (steps_done + 1) / out_difft as the reference numbers its
// files (get_data_contour receives n + temp_time + 1 and adds one more), from the pressure of
// the current time level.  The reference writes to a fixed absolute cluster path.  With
// noalloc the subsample is staged in the GPU's staging array; a zone larger than it is
// refused before anything is written.
int write_contour(lee_ctx *c, const char *dir, bool noalloc)
{
    if (!c || !dir) return set_err(c, LEE_ERR_ARG, "bad arguments");
    const Model &m = c->m;
    if (m.d.out_difft <= 0) return set_err(c, LEE_ERR_ARG, "out_difft must be positive");
    double *stage = nullptr;
    if (noalloc) {
        stage = scratch_of(c->device);
        if (!stage) return set_err(c, LEE_ERR_STATE, "no staging array on this device");
        const size_t need = std::max(contour_values(m.nx, m.ny, m.nz),
                                     m.south_porous ? contour_values(m.pnx, m.pny, m.pnz) : (size_t)0);
        if (need > kScratchDoubles)
            return set_err(c, LEE_ERR_ARG, "lee_write_contour_noalloc: a zone of " + std::to_string(need) +
                                               " values exceeds the staging array of " + std::to_string(kScratchDoubles) +
                                               " values (4 MiB per GPU); use lee_write_contour");
    }
    const long long nn = (c->steps + 1) / m.d.out_difft;
    char name[64];
    std::snprintf(name, sizeof name, "/pt%03lld.dat", nn);
    const std::string path = std::string(dir) + name;
    std::ofstream o(path.c_str(), std::ios::out | std::ios::trunc | std::ios::binary);
    if (!o) return set_err(c, LEE_ERR_IO, "cannot write " + path);
    o.setf(std::ios::scientific, std::ios::floatfield);
    o.precision(6);
    const int N = c->mf_count < 0 ? 0 : c->mf_count;
    const double ys = N * (m.trail - 1) * m.d.dy;
    const int b = c->cur * 4 + 3;
    int rc;
    if (!m.south_porous) {
        o << "VARIABLES = \"X\", \"Y\", \"Z\", \"P\"" << std::endl;
        o << "ZONE T = \"Air Zone\",";
        rc = c->prec == 64 ? contour_zone<double>(c, o, (const double *)c->air[b], m.nx, m.ny, m.nz, m.ax, m.ay, m.az, ys, stage)
                           : contour_zone<float>(c, o, (const float *)c->air[b], m.nx, m.ny, m.nz, m.ax, m.ay, m.az, ys, stage);
        if (rc) return rc;
    } else {
        // the reference omits the VARIABLES line when the ground is porous
        o << "ZONE T = \"Air Zone\",";
        rc = c->prec == 64 ? contour_zone<double>(c, o, (const double *)c->air[b], m.nx, m.ny, m.nz, m.ax, m.ay, m.az, ys, stage)
                           : contour_zone<float>(c, o, (const float *)c->air[b], m.nx, m.ny, m.nz, m.ax, m.ay, m.az, ys, stage);
        if (rc) return rc;
        const double pys = N * (m.trail - 1) * m.d.pore_dy;
        o << "ZONE T = \"south pore Zone\",";
        rc = c->prec == 64
                 ? contour_zone<double>(c, o, (const double *)c->pore[b], m.pnx, m.pny, m.pnz, m.pax, m.pay, m.paz, pys, stage)
                 : contour_zone<float>(c, o, (const float *)c->pore[b], m.pnx, m.pny, m.pnz, m.pax, m.pay, m.paz, pys, stage);
        if (rc) return rc;
    }
    if (!o) return set_err(c, LEE_ERR_IO, "write failed: " + path);
    return LEE_OK;
}

}  // namespace

extern "C" {

int lee_write_contour(lee_ctx *c, const char *dir) { return write_contour(c, dir, false); }

int lee_write_contour_noalloc(lee_ctx *c, const char *dir) { return write_contour(c, dir, true); }

int lee_get_psi(lee_ctx *c, double *psi, long long cap)
{
    if (!c || !psi) return set_err(c, LEE_ERR_ARG, "bad arguments");
    size_t n = (size_t)c->m.nx * c->m.nz * 6;
    if ((long long)n > cap) return set_err(c, LEE_ERR_ARG, "buffer too small");
    HIPCHK(c, hipMemcpyAsync(psi, c->psi, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LEE_OK;
}

int lee_get_ground_psi(lee_ctx *c, double *psi, long long cap)
{
    if (!c || !psi) return set_err(c, LEE_ERR_ARG, "bad arguments");
    if (!c->m.ground) return set_err(c, LEE_ERR_STATE, "no ground impedance in this context");
    size_t n = (size_t)c->m.nx * c->m.ny * c->m.gK;
    if ((long long)n > cap) return set_err(c, LEE_ERR_ARG, "buffer too small");
    std::vector<double> d(n);
    HIPCHK(c, hipMemcpyAsync(d.data(), c->gpsi, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ground_psi_layout(d.data(), psi, (size_t)c->m.nx * c->m.ny, c->m.gK, true);
    return LEE_OK;
}

double lee_last_step_ms(const lee_ctx *c) { return c ? (double)c->last_ms : 0.0; }

int lee_profile(lee_ctx *c, int enable)
{
    if (!c) return LEE_ERR_ARG;
    if (enable < 0 || enable > 2) return set_err(c, LEE_ERR_ARG, "profile level 0, 1 or 2");
    c->profile = enable;
    for (auto &k : c->kstat) { k.launches = 0; k.total_ms = 0; }
    return LEE_OK;
}

// Compulsory bytes of one launch: every element a kernel must read or write, once.
int lee_kernel_stats(const lee_ctx *c, lee_kstat *out, int cap)
{
    if (!c || !out) return LEE_ERR_ARG;
    const Model &m = c->m;
    const double es = (double)c->esz;
    const double inner = (double)(m.nx - 2) * (m.ny - 2) * (m.nz - 2);
    // PML auxiliaries per slab element: the pressure kernel reads Q_u, Q_v, Q_w, Q_p of x and y
    // layers (Q_u, Q_v, Q_p of z layers) and rewrites Q_p; the velocity kernel reads all four
    // and rewrites Q_u, Q_v, Q_w.
    double qp = 0, qv = 0;
    for (int s = 0; s < 6; ++s) {
        double cnt = (double)m.slab[s].count(m.nx, m.ny, m.nz);
        if (!cnt || (s == 2 && m.J3 == 0)) continue;
        bool zs = m.slab[s].axis == 2;
        qp += cnt * ((zs ? 3 : 4) + 1);
        qv += cnt * (4 + 3);
    }
    double wind = m.has_wind ? 3.0 * m.nx * m.ny * m.nz : 0.0;
    const double pinner = (double)std::max(m.pnx - 2, 0) * std::max(m.pny - 2, 0) * std::max(m.pnz - 2, 0);
    // fused pass: read u, v, w, p and write them once; per slab element read the four old
    // auxiliaries and write the four new ones
    double qf = 0;
    for (int s = 0; s < 6; ++s) {
        double cnt = (double)m.slab[s].count(m.nx, m.ny, m.nz);
        if (!cnt || (s == 2 && m.J3 == 0)) continue;
        qf += cnt * 8;
    }
    // faces: every boundary node written once from one source node (pressure) or three
    // (velocity), plus the East TDBC histories (FP64, read and written)
    const double nb = (double)boundary_count<double>(m.nx, m.ny, m.nz) +
                      (m.south_porous ? (double)boundary_count<double>(m.pnx, m.pny, m.pnz) : 0.0);
    const double psi = m.east_tdbc ? (double)m.nx * m.nz * 6 * 8 * 2 : 0.0;
    // RK4, per stage launch (a quarter of a step): fields 64 words per point and step (x_n,
    // stage state, accumulator: stages read 4/12/12/8 and write 8/8/8/4 words); FP32 reads x_n
    // and the carried state error in the last stage and writes the error back (76 words); the
    // porous velocities keep a compensated accumulator (24 words each per step, the pressure
    // 19); the auxiliaries 64 per slab element and step; the mean flow 1 (V only) or 3 words
    // per point and stage; the faces write 4 words per boundary node from 4 source values.
    const double rkw = (c->prec == 32 ? 76.0 : 64.0) / 4.0;
    const double prkw = (c->prec == 32 ? 3 * 24.0 + 19.0 : 64.0) / 4.0;
    double qrk = 0;
    for (int s = 0; s < 6; ++s) {
        double cnt = (double)m.slab[s].count(m.nx, m.ny, m.nz);
        if (!cnt || (s == 2 && m.J3 == 0)) continue;
        qrk += cnt * 16.0;
    }
    const double wrk = m.has_wind ? (m.wind_v_only ? 1.0 : 3.0) * m.nx * m.ny * m.nz : 0.0;
    // x-face nodes of the interior rows, when the RK4 air stage's launch shape takes them
    const double nxf = c->rk_xfaces ? 2.0 * (m.ny - 2) * (m.nz - 2) : 0.0;
    const double fxf = c->fb_xfaces ? 2.0 * (m.ny - 2) * (m.nz - 2) : 0.0;  // the same for the FB passes
    const char *names[KS_N] = {"air_p", "air_v", "pore_p", "pore_v", "air_fb", "bc_p", "bc_v", "rk_air", "rk_pore", "rk_bnd"};
    double bytes[KS_N] = {(5 * inner + qp + wind + 2 * fxf) * es, (7 * inner + qv + wind + 6 * fxf) * es, 5 * pinner * es,
                          7 * pinner * es, (8 * inner + qf + wind) * es, 2 * (nb - fxf) * es, 6 * (nb - fxf) * es + psi,
                          (rkw * inner + qrk + wrk + 8 * nxf) * es, prkw * pinner * es, 8 * (nb - nxf) * es};
    double pts[KS_N] = {inner, inner, pinner, pinner, inner, nb - fxf, nb - fxf, inner, pinner, nb - nxf};
    int n = 0;
    for (int k = 0; k < KS_N && n < cap; ++k) {
        lee_kstat s = c->kstat[k];
        std::snprintf(s.name, sizeof s.name, "%s", names[k]);
        s.bytes_per_launch = bytes[k];
        s.points_per_launch = pts[k];
        out[n++] = s;
    }
    return n;
}

double lee_copy_bench(lee_ctx *c, long long n_words, int repeats)
{
    if (!c || n_words <= 0 || repeats <= 0) return -1.0;
    void *a = nullptr, *b = nullptr;
    size_t bytes = (size_t)n_words * 4;
    if (hipMalloc(&a, bytes) != hipSuccess) return -1.0;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -1.0; }
    (void)hipMemsetAsync(a, 0, bytes, c->stream);
    launch_copy_words((float *)b, (const float *)a, (size_t)n_words, c->stream);  // warm-up
    (void)hipEventRecord(c->ev0, c->stream);
    for (int r = 0; r < repeats; ++r) launch_copy_words((float *)b, (const float *)a, (size_t)n_words, c->stream);
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    (void)hipFree(a);
    (void)hipFree(b);
    return (double)ms / repeats;
}

double lee_stream_bench(lee_ctx *c, long long n_words, int reads, int writes, int repeats)
{
    if (!c || n_words <= 0 || repeats <= 0 || reads < 1 || reads > kStreamMaxR || writes < 1 || writes > kStreamMaxW)
        return -1.0;
    std::vector<float *> arr;
    auto release = [&]() { for (float *p : arr) (void)hipFree(p); };
    const size_t bytes = (size_t)n_words * 4;
    for (int q = 0; q < reads + writes; ++q) {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) { release(); return -1.0; }
        arr.push_back((float *)p);
        (void)hipMemsetAsync(p, 0, bytes, c->stream);
    }
    const float *const *in = arr.data();
    float *const *out = arr.data() + reads;
    if (!launch_stream_words(out, writes, in, reads, (size_t)n_words, c->stream)) { release(); return -1.0; }  // warm-up
    (void)hipEventRecord(c->ev0, c->stream);
    for (int r = 0; r < repeats; ++r) (void)launch_stream_words(out, writes, in, reads, (size_t)n_words, c->stream);
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0.f;
    const bool ok = hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess && hipGetLastError() == hipSuccess;
    release();
    return ok ? (double)ms / repeats : -1.0;
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// Ensemble: many independent domains advanced together.  Every RK4 stage is three launches
// for all members (air, porous layer, faces); each block finds its member through a block
// prefix and reads that member's kernel arguments from a device table.  Members are full
// contexts (state, receivers, restart files) that lee_ensemble_member exposes.
// ------------------------------------------------------------------------------------
struct lee_ensemble {
    std::vector<lee_ctx *> m;
    int prec = 32, wind = 0;
    bool terrain = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    std::string err;
    std::vector<void *> allocs;
    void *air_args[4] = {}, *pore_args[4] = {}, *bnd_args[4] = {};
    // tile prefixes [air n+1 | pore np+1 | bnd n+1 | air n+1 | pore np+1] (np porous
    // members): the last two with each member plane's tiles padded to a multiple of 8
    // (launch form 1)
    int *first = nullptr;
    int air_blocks = 0, pore_blocks = 0, bnd_blocks = 0, n_pore = 0, air_blocks_pad = 0, pore_blocks_pad = 0;
    int form = -1;       // launch form of the air stage (kEnsForms), -1 until tuned
    int form_pore = -1;  // and of the porous stage (tuned on its own)
    int form_pin = 0;    // lee_ensemble_set_launch: 0 autotune, 1 + form pinned for both
    std::vector<int> key;  // (J1, y1on) per member the tables hold
    int version = 0;
    void **rec_p = nullptr;
    long long *rec_idx = nullptr;
    double *rec_buf = nullptr;
    int nrec = 0;
    std::vector<int> rec_off;
    struct Graph { int n, version; hipGraphExec_t exec; };
    std::vector<Graph> graphs;

    ~lee_ensemble()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        for (void *p : allocs) (void)hipFree(p);
        for (lee_ctx *c : m) delete c;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int ens_err(lee_ensemble *e, int code, const std::string &msg)
{
    if (e) e->err = msg;
    set_global_error(msg);
    return code;
}

#define EHIP(e, expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ens_err((e), LEE_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ens_alloc(lee_ensemble *e, void **out, size_t bytes)
{
    void *p = nullptr;
    EHIP(e, hipMalloc(&p, std::max<size_t>(bytes, 8)));
    e->allocs.push_back(p);
    *out = p;
    return LEE_OK;
}

// Argument tables of the four stages for the members' current gate states.
template <typename T> int ens_tables(lee_ensemble *e, const std::vector<int> &J1, const std::vector<int> &y1on)
{
    const int n = (int)e->m.size();
    EHIP(e, hipStreamSynchronize(e->stream));  // no launch may still read the old tables
    for (int st = 0; st < 4; ++st) {
        std::vector<RKArgs<T>> ra(n);
        std::vector<RKPoreArgs<T>> pa;
        std::vector<RKBndArgs<T>> ba(n);
        for (int d = 0; d < n; ++d) {
            RKPoreArgs<T> pr;
            rk4_stage_args<T>(e->m[d], st, J1[d], y1on[d], ra[d], pr, ba[d]);
            // the batched air stage takes the x-face nodes whose neighbour is in its wavefront
            ba[d].b.xring = 2;
            if (e->m[d]->m.south_porous) pa.push_back(pr);
        }
        EHIP(e, hipMemcpy(e->air_args[st], ra.data(), n * sizeof(RKArgs<T>), hipMemcpyHostToDevice));
        EHIP(e, hipMemcpy(e->bnd_args[st], ba.data(), n * sizeof(RKBndArgs<T>), hipMemcpyHostToDevice));
        if (!pa.empty())
            EHIP(e, hipMemcpy(e->pore_args[st], pa.data(), pa.size() * sizeof(RKPoreArgs<T>), hipMemcpyHostToDevice));
    }
    for (int d = 0; d < n; ++d) { e->key[2 * d] = J1[d]; e->key[2 * d + 1] = y1on[d]; }
    e->version += 1;
    return LEE_OK;
}

// The air launch (form fa, -1 none) and the porous launch (form fp, -1 none) of RK4 stage st.
template <typename T> void ens_stage_cells(lee_ensemble *e, int st, int fa, int fp, hipStream_t s)
{
    const int n = (int)e->m.size();
    const int *pa = e->first + (fa == 1 ? 2 * (n + 1) + e->n_pore + 1 : 0);
    const int *pp = e->first + (fp == 1 ? 3 * (n + 1) + e->n_pore + 1 : n + 1);
    if (fa >= 0)
        launch_rk_air_batch<T>((const RKArgs<T> *)e->air_args[st], pa, n, fa == 1 ? e->air_blocks_pad : e->air_blocks,
                               e->wind, e->terrain, fa, s);
    if (fp >= 0 && e->n_pore)
        launch_rk_pore_batch<T>((const RKPoreArgs<T> *)e->pore_args[st], pp, e->n_pore,
                                fp == 1 ? e->pore_blocks_pad : e->pore_blocks, e->terrain, fp, s);
}

// Tables current for the coming step (its gate states peeked, the counters left as they are).
template <typename T> int ens_tables_peek(lee_ensemble *e)
{
    std::vector<int> J1(e->m.size()), y1on(e->m.size());
    bool stale = false;
    for (size_t d = 0; d < e->m.size(); ++d) {
        lee_ctx *c = e->m[d];
        const long long save = c->counter;
        gate_tick(c, J1[d], y1on[d]);
        c->counter = save;
        stale |= e->key[2 * d] != J1[d] || e->key[2 * d + 1] != y1on[d];
    }
    return stale ? ens_tables<T>(e, J1, y1on) : LEE_OK;
}

// Autotuner of the launch forms: times stage 0's air launch, then its porous launch, of the
// coming step in each form (stage 0 reads only level n, so repeating it changes nothing) and
// keeps the fastest for each kind; every form gives the same bits.
template <typename T> int ens_tune(lee_ensemble *e)
{
    if (e->form_pin > 0) {
        e->form = e->form_pore = e->form_pin - 1;
        return LEE_OK;
    }
    int rc = ens_tables_peek<T>(e);
    if (rc) return rc;
    int pick[2] = {0, 0};
    for (int kind = 0; kind < (e->n_pore ? 2 : 1); ++kind) {
        float t[kEnsForms];
        for (int f = 0; f < kEnsForms; ++f) t[f] = 1e30f;
        for (int pass = 0; pass < 2; ++pass)  // each form's best of two passes, as tune_rk
            for (int f = 0; f < kEnsForms; ++f) {
                const int fa = kind == 0 ? f : -1, fp = kind == 1 ? f : -1;
                ens_stage_cells<T>(e, 0, fa, fp, e->stream);  // warm-up
                EHIP(e, hipEventRecord(e->ev0, e->stream));
                for (int rep = 0; rep < 4; ++rep) ens_stage_cells<T>(e, 0, fa, fp, e->stream);
                EHIP(e, hipEventRecord(e->ev1, e->stream));
                EHIP(e, hipEventSynchronize(e->ev1));
                float ms = 0.f;
                EHIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
                t[f] = std::min(t[f], ms);
            }
        for (int f = 1; f < kEnsForms; ++f)
            if (t[f] < t[pick[kind]]) pick[kind] = f;
    }
    EHIP(e, hipGetLastError());
    e->form = pick[0];
    e->form_pore = pick[1];
    return LEE_OK;
}

// One ensemble RK4 step (launches only once the tables are current).
template <typename T> int ens_step(lee_ensemble *e, double *rec_slot)
{
    const int n = (int)e->m.size();
    std::vector<int> J1(n), y1on(n);
    bool stale = false;
    for (int d = 0; d < n; ++d) {
        gate_tick(e->m[d], J1[d], y1on[d]);
        stale |= e->key[2 * d] != J1[d] || e->key[2 * d + 1] != y1on[d];
    }
    if (stale) {
        int rc = ens_tables<T>(e, J1, y1on);
        if (rc) return rc;
    }
    hipStream_t s = e->stream;
    launch_sample_batch<T>((const T *const *)e->rec_p, e->rec_idx, e->nrec, rec_slot, s);
    const int *fb = e->first + n + 1 + e->n_pore + 1;
    for (int st = 0; st < 4; ++st) {
        ens_stage_cells<T>(e, st, e->form, e->form_pore, s);
        launch_rk_bnd_batch<T>((const RKBndArgs<T> *)e->bnd_args[st], fb, n, e->bnd_blocks, s);
    }
    for (lee_ctx *c : e->m) c->steps += 1;
    return LEE_OK;
}

// Steps from a captured graph when no member's gate changes inside the chunk.
bool ens_gate_constant(const lee_ensemble *e, int n)
{
    for (const lee_ctx *c : e->m)
        if (!gate_constant(c, n)) return false;
    return true;
}

template <typename T> int ens_run(lee_ensemble *e, int nsteps, double *rec_out, int use_graph)
{
    hipStream_t s = e->stream;
    if (e->form < 0) {
        int rc = ens_tune<T>(e);
        if (rc) return rc;
    }
    EHIP(e, hipEventRecord(e->ev0, s));
    int done = 0;
    while (done < nsteps) {
        int n = std::min(nsteps - done, kGraphSteps);
        bool replayed = false;
        if (use_graph && ens_gate_constant(e, n)) {
            // the gate states hold for the whole chunk: bring the tables up to date before
            // capturing
            int rc = ens_tables_peek<T>(e);
            if (rc) return rc;
            hipGraphExec_t exec = nullptr;
            for (auto &g : e->graphs)
                if (g.n == n && g.version == e->version) exec = g.exec;
            if (exec) {
                EHIP(e, hipGraphLaunch(exec, s));
                for (lee_ctx *c : e->m) {
                    const Model &m = c->m;
                    if (m.judge[1] == 1 || m.judge[1] == 3) c->counter += n;
                    c->steps += n;
                }
            } else {
                EHIP(e, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int t = 0; t < n; ++t) {
                    int rc = ens_step<T>(e, e->rec_buf + (size_t)t * e->nrec);
                    if (rc) { hipGraph_t g0; (void)hipStreamEndCapture(s, &g0); return rc; }
                }
                hipGraph_t graph = nullptr;
                EHIP(e, hipStreamEndCapture(s, &graph));
                lee_ensemble::Graph g{n, e->version, nullptr};
                hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                EHIP(e, ie);
                if (e->graphs.size() >= 4) {
                    (void)hipGraphExecDestroy(e->graphs.front().exec);
                    e->graphs.erase(e->graphs.begin());
                }
                e->graphs.push_back(g);
                EHIP(e, hipGraphLaunch(g.exec, s));
            }
            replayed = true;
        }
        if (!replayed) {
            n = 1;
            int rc = ens_step<T>(e, e->rec_buf);
            if (rc) return rc;
        }
        if (rec_out && e->nrec) {
            EHIP(e, hipMemcpyAsync(rec_out + (size_t)done * e->nrec, e->rec_buf, (size_t)n * e->nrec * 8,
                                   hipMemcpyDeviceToHost, s));
            EHIP(e, hipStreamSynchronize(s));
        }
        done += n;
    }
    EHIP(e, hipEventRecord(e->ev1, s));
    EHIP(e, hipEventSynchronize(e->ev1));
    EHIP(e, hipEventElapsedTime(&e->last_ms, e->ev0, e->ev1));
    return LEE_OK;
}

}  // namespace

extern "C" {

int lee_ensemble_create(int n, const lee_deck *decks, const lee_options *opts, lee_ensemble **out)
{
    if (!out || n < 1 || !decks) return ens_err(nullptr, LEE_ERR_ARG, "bad arguments");
    *out = nullptr;
    std::unique_ptr<lee_ensemble> e(new lee_ensemble());
    lee_ensemble *E = e.get();
    for (int d = 0; d < n; ++d) {
        lee_options o;
        lee_options_default(&o);
        if (opts) o = opts[d];
        lee_ctx *c = nullptr;
        int rc = create_impl(&decks[d], &o, &c);
        if (rc) return ens_err(E, rc, "member " + std::to_string(d) + ": " + g_global_err);
        E->m.push_back(c);
        const Model &m = c->m;
        if (!m.rk4) return ens_err(E, LEE_ERR_UNSUPPORTED, "ensemble members step with RK4 (integrator = LEE_INT_RK4)");
        if (m.d.mf_judge == 1) return ens_err(E, LEE_ERR_UNSUPPORTED, "ensemble members have a fixed frame");
        if (d == 0) {
            E->prec = c->prec; E->wind = rk_wind_mode(m); E->terrain = m.terrain;
        } else if (c->prec != E->prec || rk_wind_mode(m) != E->wind || m.terrain != E->terrain) {
            return ens_err(E, LEE_ERR_UNSUPPORTED, "ensemble members share precision, wind kind and grid kind");
        }
        rc = lee_frame_setup(c, 0);
        if (rc) return ens_err(E, rc, "member " + std::to_string(d) + ": " + c->err);
        if (c->src_ps.size()) return ens_err(E, LEE_ERR_UNSUPPORTED, "no time-series source in an ensemble");
    }
    EHIP(E, hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    EHIP(E, hipEventCreate(&E->ev0));
    EHIP(E, hipEventCreate(&E->ev1));
    int rc;
    for (int st = 0; st < 4; ++st) {
        const size_t sa = E->prec == 64 ? sizeof(RKArgs<double>) : sizeof(RKArgs<float>);
        const size_t sp = E->prec == 64 ? sizeof(RKPoreArgs<double>) : sizeof(RKPoreArgs<float>);
        const size_t sb = E->prec == 64 ? sizeof(RKBndArgs<double>) : sizeof(RKBndArgs<float>);
        if ((rc = ens_alloc(E, &E->air_args[st], n * sa)) || (rc = ens_alloc(E, &E->pore_args[st], n * sp)) ||
            (rc = ens_alloc(E, &E->bnd_args[st], n * sb)))
            return rc;
    }
    // tile prefixes: interior columns of the air and porous grids in 256s per plane (and the
    // same with each plane's tiles padded to a multiple of 8), boundary nodes in 256s
    std::vector<int> fa(1, 0), fp(1, 0), fb(1, 0), fa8(1, 0), fp8(1, 0);
    for (lee_ctx *c : E->m) {
        const Model &m = c->m;
        long long ta = batch_blocks(m.nx, m.ny, m.nz), ta8 = batch_blocks(m.nx, m.ny, m.nz, true);
        long long nb = (c->prec == 64 ? boundary_count<double>(m.nx, m.ny, m.nz) : boundary_count<float>(m.nx, m.ny, m.nz));
        if (m.south_porous) {
            fp.push_back((int)(fp.back() + batch_blocks(m.pnx, m.pny, m.pnz)));
            fp8.push_back((int)(fp8.back() + batch_blocks(m.pnx, m.pny, m.pnz, true)));
            nb += c->prec == 64 ? boundary_count<double>(m.pnx, m.pny, m.pnz) : boundary_count<float>(m.pnx, m.pny, m.pnz);
        }
        if (fa8.back() + ta8 > INT_MAX / 2 || fp8.back() > INT_MAX / 2 || fb.back() + (nb + 255) / 256 > INT_MAX / 2)
            return ens_err(E, LEE_ERR_ARG, "ensemble too large for one launch");
        fa.push_back((int)(fa.back() + ta));
        fa8.push_back((int)(fa8.back() + ta8));
        fb.push_back((int)(fb.back() + (nb + 255) / 256));
    }
    E->air_blocks = fa.back(); E->pore_blocks = fp.back(); E->bnd_blocks = fb.back();
    E->air_blocks_pad = fa8.back(); E->pore_blocks_pad = fp8.back();
    E->n_pore = (int)fp.size() - 1;
    std::vector<int> first(fa);
    first.insert(first.end(), fp.begin(), fp.end());
    first.insert(first.end(), fb.begin(), fb.end());
    first.insert(first.end(), fa8.begin(), fa8.end());
    first.insert(first.end(), fp8.begin(), fp8.end());
    if ((rc = ens_alloc(E, (void **)&E->first, first.size() * sizeof(int)))) return rc;
    EHIP(E, hipMemcpy(E->first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
    // receivers of all members, concatenated
    std::vector<void *> rp;
    std::vector<long long> ri;
    for (lee_ctx *c : E->m) {
        E->rec_off.push_back((int)ri.size());
        const Model &m = c->m;
        for (int r = 0; r < (int)m.ri.size(); ++r) {
            rp.push_back(c->air[c->cur * 4 + 3]);
            ri.push_back(m.receiver_index(r, 0));
        }
    }
    E->rec_off.push_back((int)ri.size());
    E->nrec = (int)ri.size();
    if ((rc = ens_alloc(E, (void **)&E->rec_p, rp.size() * sizeof(void *))) ||
        (rc = ens_alloc(E, (void **)&E->rec_idx, ri.size() * sizeof(long long))) ||
        (rc = ens_alloc(E, (void **)&E->rec_buf, (size_t)kGraphSteps * std::max(E->nrec, 1) * 8)))
        return rc;
    if (E->nrec) {
        EHIP(E, hipMemcpy(E->rec_p, rp.data(), rp.size() * sizeof(void *), hipMemcpyHostToDevice));
        EHIP(E, hipMemcpy(E->rec_idx, ri.data(), ri.size() * sizeof(long long), hipMemcpyHostToDevice));
    }
    E->key.assign(2 * n, -1);  // forces the first table build
    *out = e.release();
    return LEE_OK;
}

int lee_ensemble_destroy(lee_ensemble *e)
{
    delete e;
    return LEE_OK;
}

int lee_ensemble_size(const lee_ensemble *e) { return e ? (int)e->m.size() : 0; }

int lee_ensemble_member(lee_ensemble *e, int d, lee_ctx **ctx)
{
    if (!e || !ctx || d < 0 || d >= (int)e->m.size()) return ens_err(e, LEE_ERR_ARG, "bad member");
    *ctx = e->m[d];
    return LEE_OK;
}

int lee_ensemble_receivers(const lee_ensemble *e, int *offsets, int cap)
{
    if (!e) return 0;
    const int n = (int)e->rec_off.size();
    for (int d = 0; d < n && d < cap; ++d) offsets[d] = e->rec_off[d];
    return n;
}

int lee_ensemble_step(lee_ensemble *e, int nsteps, double *rec_out, int use_graph)
{
    if (!e || nsteps < 0) return ens_err(e, LEE_ERR_ARG, "bad arguments");
    if (nsteps == 0) return LEE_OK;
    return e->prec == 64 ? ens_run<double>(e, nsteps, rec_out, use_graph) : ens_run<float>(e, nsteps, rec_out, use_graph);
}

double lee_ensemble_last_ms(const lee_ensemble *e) { return e ? (double)e->last_ms : 0.0; }

int lee_ensemble_set_launch(lee_ensemble *e, int form)
{
    if (!e || form < 0 || form > kEnsForms) return ens_err(e, LEE_ERR_ARG, "launch form 0 (autotune) .. 3");
    if (e->stream) EHIP(e, hipStreamSynchronize(e->stream));
    for (auto &g : e->graphs) (void)hipGraphExecDestroy(g.exec);  // captured with the old form
    e->graphs.clear();
    e->form_pin = form;
    e->form = e->form_pore = -1;
    return LEE_OK;
}

int lee_ensemble_launch(const lee_ensemble *e, int porous)
{
    if (!e) return 0;
    const int f = porous ? e->form_pore : e->form;
    return f >= 0 ? f + 1 : 0;
}

const char *lee_ensemble_error(const lee_ensemble *e) { return e ? e->err.c_str() : g_global_err.c_str(); }

}  // extern "C"
