/*
 * ref_probe.cpp -- drives the reference solver's own compiled stencil code.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference sources where
 * they lie (air.cpp, porous.cpp, df_over_var.cpp, decomp.cpp, FFT.cpp under
 * 3D_MPI_TEST/) into oracle/_ref/ref_probe; nothing of the reference is copied here.
 *
 * calculation_2D.cpp and airpore.cpp need the MSVC CRT (strcat_s/sprintf_s array
 * templates, <direct.h>) and are not built, so the objects are never constructed: the
 * probe lays an `air` / `porous` object out in zeroed storage, fills the members that
 * calculation_2D's constructor and SetMovingFrame would set (calculation_2D.cpp:12-236,
 * 405-433 -- restated below), and then calls the reference's own member functions:
 *   air::cal_PML_var (air.cpp:509-937), SpeedProfile (air.cpp:40-256),
 *   air::SetWindProfile (air.cpp:1286-1304), air::SetQMove (air.cpp:2447-2730),
 *   air::cal_fp (air.cpp:1953-2322), air::cal_fuvw (air.cpp:1476-1949),
 *   air::Update_PML_Qp / Qvw (air.cpp:2323-2439),
 *   porous::cal_fp / cal_fuvw (porous.cpp:77-143), FFT / FFT_output (FFT.cpp:15-121),
 *   air::cal_hill_var (air.cpp:938-1105),
 *   Geo_over_var (df_over_var.cpp:5-9), mpe_decomp1d (decomp.cpp:4-19).
 * Calls are qualified (non-virtual), so no vtable is needed.  MPI runs as a singleton.
 *
 * Protocol: argv[1] = case file, argv[2] = result file.  Case file: little-endian
 * records  int32 count + payload, in the order read by main().  Result file: records
 * [int32 name_len][name][int64 count][count doubles].
 * Field arrays in both files use the x-fastest layout idx = i + nx*(j + ny*k).
 */
#include "air.h"
#include "porous.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace {

FILE *g_out = nullptr;

void emit(const char *name, const std::vector<double> &v)
{
    int32_t n = (int32_t)strlen(name);
    int64_t c = (int64_t)v.size();
    fwrite(&n, 4, 1, g_out);
    fwrite(name, 1, n, g_out);
    fwrite(&c, 8, 1, g_out);
    if (c) fwrite(v.data(), 8, c, g_out);
}

struct Reader {
    FILE *f;
    std::vector<double> d()
    {
        int32_t n = 0;
        if (fread(&n, 4, 1, f) != 1) { fprintf(stderr, "short read\n"); exit(3); }
        std::vector<double> v(n);
        if (n && fread(v.data(), 8, n, f) != (size_t)n) { fprintf(stderr, "short read\n"); exit(3); }
        return v;
    }
};

double ***tri(int a, int b, int c)
{
    double ***t = new double **[a];
    for (int i = 0; i < a; ++i) {
        t[i] = new double *[b];
        for (int j = 0; j < b; ++j) t[i][j] = new double[c]();
    }
    return t;
}

/* deterministic fill for PML auxiliaries: values of size ~1e-3 with no structure */
double hashval(unsigned long long x)
{
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
    return ((double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 2e-3;
}

struct Grid {
    int nx, ny, nz;
    std::vector<double> X, Y, Z;
};

struct AirProbe : public air {
    void setup(const Grid &g, const int *hdr, double dt, double dx, double dy, double dz,
               double gam, double Pav_, double c, int vmethod, double vcoef,
               int jmf, int lead, int trail, MPI_Comm comm)
    {
        IMAX = hdr[0]; IMAX1 = hdr[1]; IMAX2 = hdr[2]; JMAX = g.ny; JMAX1 = hdr[4]; JMAX2 = hdr[5];
        JMAX3 = JMAX1; KMAX = hdr[6]; KMAX1 = hdr[7]; KMAX2 = hdr[8]; IJKMAX = hdr[9];
        diff_t = dt; diff_x = 1.0 / dx; diff_y = 1.0 / dy; diff_z = 1.0 / dz;
        diff_x1 = dx; diff_y1 = dy; diff_z1 = dz;
        scheme = FB_p_v;
        move_frame.Judge = 1; move_frame.JMAX = jmf; move_frame.lead_DJ = lead; move_frame.trail_DJ = trail;
        move_frame.ratio = 1.0;
        mpi_rank = 0; mpi_size = 1; mpi_xarea = mpi_yarea = mpi_zarea = 1;
        mpi_coords = new int[3](); mpi_comm3d = comm;
        MPI_Cart_shift(comm, 0, 1, &mpi_nback, &mpi_nfront);
        MPI_Cart_shift(comm, 1, 1, &mpi_nleft, &mpi_nright);
        MPI_Cart_shift(comm, 2, 1, &mpi_nbottom, &mpi_ntop);
        mpi_i1 = 0; mpi_i2 = g.nx - 1; mpi_j1 = 0; mpi_j2 = g.ny - 1; mpi_k1 = 0; mpi_k2 = g.nz - 1;
        mpi_IMAX = g.nx; mpi_JMAX = g.ny; mpi_KMAX = g.nz;
        X = new double[g.nx]; Y = new double[g.ny]; Z = new double[g.nz];
        axisbound_X = new double[g.nx]; axisbound_Y = new double[g.ny]; axisbound_Z = new double[g.nz];
        for (int i = 0; i < g.nx; ++i) X[i] = axisbound_X[i] = g.X[i];
        for (int j = 0; j < g.ny; ++j) Y[j] = axisbound_Y[j] = g.Y[j];
        for (int k = 0; k < g.nz; ++k) Z[k] = axisbound_Z[k] = g.Z[k];
        X_over_x = new double[g.nx](); Y_over_y = new double[g.ny](); Z_over_z = new double[g.nz]();
        X_over_y = X_over_z = Y_over_x = Y_over_z = Z_over_x = Z_over_y = 0.0;
        Jacobian = tri(g.nx, g.ny, g.nz);
        u_n = tri(g.nx, g.ny, g.nz); v_n = tri(g.nx, g.ny, g.nz); w_n = tri(g.nx, g.ny, g.nz); p_n = tri(g.nx, g.ny, g.nz);
        u_nn = tri(g.nx, g.ny, g.nz); v_nn = tri(g.nx, g.ny, g.nz); w_nn = tri(g.nx, g.ny, g.nz); p_nn = tri(g.nx, g.ny, g.nz);
        int NN = lead * g.nx * g.nz;
        uer = new double[NN](); uws = new double[NN](); ver = new double[NN](); vws = new double[NN]();
        wer = new double[NN](); wws = new double[NN](); per = new double[NN](); pws = new double[NN]();
        adiabatic_coef = gam; Pav = Pav_; sound_speed = c;
        Aav = pow(c, 2) / gam / Pav_; /* airpore.cpp:177 */
        velocity_method = vmethod;
        SP = new SpeedProfile(vcoef, c, vmethod);
        Uav = tri(g.nx, g.ny, g.nz); Vav = tri(g.nx, g.ny, g.nz); Wav = tri(g.nx, g.ny, g.nz);
        cr_judge = 0;
    }
    /* metric part of calculation_2D::SetMovingFrame (calculation_2D.cpp:410-433) */
    void metrics(int N)
    {
        if (N != 0)
            for (int j = 0; j < mpi_JMAX; ++j) Y[j] = axisbound_Y[mpi_j1 + j] + N * (move_frame.trail_DJ - 1) * diff_y1;
        for (int i = 1; i < mpi_IMAX - 1; i++) X_over_x[i] = (X[i + 1] - X[i - 1]) * diff_x * 0.5;
        for (int j = 1; j < mpi_JMAX - 1; j++) Y_over_y[j] = (Y[j + 1] - Y[j - 1]) * diff_y * 0.5;
        for (int k = 1; k < mpi_KMAX - 1; k++) Z_over_z[k] = (Z[k + 1] - Z[k - 1]) * diff_z * 0.5;
        for (int i = 1; i < mpi_IMAX - 1; i++)
            for (int j = 1; j < mpi_JMAX - 1; j++)
                for (int k = 1; k < mpi_KMAX - 1; k++) {
                    double J = X_over_x[i] * Y_over_y[j] * Z_over_z[k] + Z_over_x * X_over_y * Y_over_z + Y_over_x * Z_over_y * X_over_z;
                    J += -X_over_x[i] * Z_over_y * Y_over_z - Y_over_x * X_over_y * Z_over_z[k] - Z_over_x * Y_over_y[j] * X_over_z;
                    Jacobian[i][j][k] = 1.0 / J;
                }
    }
    void load(double ***t, const std::vector<double> &f)
    {
        for (int k = 0; k < mpi_KMAX; ++k)
            for (int j = 0; j < mpi_JMAX; ++j)
                for (int i = 0; i < mpi_IMAX; ++i) t[i][j][k] = f[i + (size_t)mpi_IMAX * (j + (size_t)mpi_JMAX * k)];
    }
    std::vector<double> dump(double ***t)
    {
        std::vector<double> f((size_t)mpi_IMAX * mpi_JMAX * mpi_KMAX, 0.0);
        for (int k = 0; k < mpi_KMAX; ++k)
            for (int j = 0; j < mpi_JMAX; ++j)
                for (int i = 0; i < mpi_IMAX; ++i) f[i + (size_t)mpi_IMAX * (j + (size_t)mpi_JMAX * k)] = t[i][j][k];
        return f;
    }
    /* Q slabs in x-fastest layout of each slab: x slab [ii][j][k] -> ii + n*(j + ny*k), ... */
    std::vector<double> qdump(double ***q, int axis, int n)
    {
        std::vector<double> f;
        if (!n) return f;
        int a = axis == 0 ? n : mpi_IMAX, b = axis == 1 ? n : mpi_JMAX, c = axis == 2 ? n : mpi_KMAX;
        f.assign((size_t)a * b * c, 0.0);
        for (int k = 0; k < c; ++k)
            for (int j = 0; j < b; ++j)
                for (int i = 0; i < a; ++i) f[i + (size_t)a * (j + (size_t)b * k)] = q[i][j][k];
        return f;
    }
    void qfill(double ***q, int axis, int n, unsigned long long salt)
    {
        if (!n) return;
        int a = axis == 0 ? n : mpi_IMAX, b = axis == 1 ? n : mpi_JMAX, c = axis == 2 ? n : mpi_KMAX;
        for (int i = 0; i < a; ++i)
            for (int j = 0; j < b; ++j)
                for (int k = 0; k < c; ++k) q[i][j][k] = hashval(salt * 1000003ULL + (unsigned long long)(i + a * (j + (unsigned long long)b * k)));
    }
    struct Slab { double ****q; int axis; int *n; const char *name; };
    void slabs(Slab *s)
    {
        static double ***dummy = nullptr;
        (void)dummy;
        s[0] = {nullptr, 0, &mpi_IMAX1, "x1"}; s[1] = {nullptr, 0, &mpi_IMAX2, "x2"};
        s[2] = {nullptr, 1, &mpi_JMAX1, "y1"}; s[3] = {nullptr, 1, &mpi_JMAX2, "y2"};
        s[4] = {nullptr, 2, &mpi_KMAX1, "z1"}; s[5] = {nullptr, 2, &mpi_KMAX2, "z2"};
    }
    double ***&Q(int slab, int f)
    {
        double ****tab[6][4] = {{&Qux1, &Qvx1, &Qwx1, &Qpx1}, {&Qux2, &Qvx2, &Qwx2, &Qpx2},
                                {&Quy1, &Qvy1, &Qwy1, &Qpy1}, {&Quy2, &Qvy2, &Qwy2, &Qpy2},
                                {&Quz1, &Qvz1, &Qwz1, &Qpz1}, {&Quz2, &Qvz2, &Qwz2, &Qpz2}};
        return *tab[slab][f];
    }
    int qn(int slab)
    {
        int v[6] = {mpi_IMAX1, mpi_IMAX2, mpi_JMAX1, mpi_JMAX2, mpi_KMAX1, mpi_KMAX2};
        return v[slab];
    }
    void emit_q(const char *tag)
    {
        const char *sn[6] = {"x1", "x2", "y1", "y2", "z1", "z2"};
        const char *fn[4] = {"u", "v", "w", "p"};
        int axis[6] = {0, 0, 1, 1, 2, 2};
        for (int s = 0; s < 6; ++s)
            for (int f = 0; f < 4; ++f) {
                std::string nm = std::string(tag) + "_" + sn[s] + "_" + fn[f];
                emit(nm.c_str(), qdump(Q(s, f), axis[s], qn(s)));
            }
    }
    void fill_q(unsigned long long seed)
    {
        int axis[6] = {0, 0, 1, 1, 2, 2};
        for (int s = 0; s < 6; ++s)
            for (int f = 0; f < 4; ++f) qfill(Q(s, f), axis[s], qn(s), seed + 17ULL * s + f);
    }
    int gateJ1() const { return JMAX1; }
    void set_gate(int counter, int J1) { PML_JMAX1 = counter; JMAX1 = J1; }
    void pml(PML_boundary x, PML_boundary y, PML_boundary z) { air::cal_PML_var(x, y, z); }
    void wind(int N) { air::SetWindProfile(N); }
    void qmove(int N) { air::SetQMove(N); }
    void fp() { air::cal_fp(u_n, v_n, w_n, p_n); }
    void fuvw() { air::cal_fuvw(u_n, v_n, w_n, p_n); }
    void qp() { air::Update_PML_Qp(); }
    void qvw() { air::Update_PML_Qvw(); }
    std::vector<double> U() { return dump(Uav); }
    std::vector<double> V() { return dump(Vav); }
    std::vector<double> W() { return dump(Wav); }
    double ***un() { return u_n; } double ***vn() { return v_n; } double ***wn() { return w_n; } double ***pn() { return p_n; }
    double ***unn() { return u_nn; } double ***vnn() { return v_nn; } double ***wnn() { return w_nn; } double ***pnn() { return p_nn; }
    /* embedded obstacle: air::cal_hill_var (air.cpp:938-1105), then its box and mask */
    void hill(hillpore h, hillpore c, hillpore t) { air::cal_hill_var(h, c, t); }
    std::vector<double> hill_box() const
    {
        return {(double)istar, (double)istop, (double)jstar, (double)jstop, (double)kstar, (double)kstop};
    }
    std::vector<double> hill_mask() const
    {
        std::vector<double> m;
        for (int k = kstar; k <= kstop; ++k)
            for (int j = jstar; j <= jstop; ++j)
                for (int i = istar; i <= istop; ++i) m.push_back(cr_uc[i - istar][j - jstar][k - kstar]);
        return m;
    }
};

struct PoreProbe : public porous {
    void setup(const Grid &g, double dt, double dx, double dy, double dz, double gam, double Pav_,
               double c, double q, double poro, double sig)
    {
        diff_t = dt; diff_x = 1.0 / dx; diff_y = 1.0 / dy; diff_z = 1.0 / dz;
        diff_x1 = dx; diff_y1 = dy; diff_z1 = dz;
        mpi_IMAX = g.nx; mpi_JMAX = g.ny; mpi_KMAX = g.nz; mpi_j1 = 0;
        X = new double[g.nx]; Y = new double[g.ny]; Z = new double[g.nz];
        for (int i = 0; i < g.nx; ++i) X[i] = g.X[i];
        for (int j = 0; j < g.ny; ++j) Y[j] = g.Y[j];
        for (int k = 0; k < g.nz; ++k) Z[k] = g.Z[k];
        X_over_x = new double[g.nx](); Y_over_y = new double[g.ny](); Z_over_z = new double[g.nz]();
        X_over_y = X_over_z = Y_over_x = Y_over_z = Z_over_x = Z_over_y = 0.0;
        Jacobian = tri(g.nx, g.ny, g.nz);
        u_n = tri(g.nx, g.ny, g.nz); v_n = tri(g.nx, g.ny, g.nz); w_n = tri(g.nx, g.ny, g.nz); p_n = tri(g.nx, g.ny, g.nz);
        u_nn = tri(g.nx, g.ny, g.nz); v_nn = tri(g.nx, g.ny, g.nz); w_nn = tri(g.nx, g.ny, g.nz); p_nn = tri(g.nx, g.ny, g.nz);
        /* PorePara derivation (airpore.cpp:249-251) then porous::porous (porous.cpp:40-51) */
        double Aav_ = pow(c, 2) / gam / Pav_;
        double effd = q * q / poro / Aav_;
        double Kp_ = 1.0 / gam / Pav_;
        eff_density = 1.0 / effd; porosity = poro; resistivity = sig; Kp = Kp_;
        Beta = resistivity * diff_t * eff_density;
        Gama = diff_t / Kp / porosity;
        invBeta = 1.0 / (Beta + 1.0);
        for (int i = 1; i < mpi_IMAX - 1; i++) X_over_x[i] = (X[i + 1] - X[i - 1]) * diff_x * 0.5;
        for (int j = 1; j < mpi_JMAX - 1; j++) Y_over_y[j] = (Y[j + 1] - Y[j - 1]) * diff_y * 0.5;
        for (int k = 1; k < mpi_KMAX - 1; k++) Z_over_z[k] = (Z[k + 1] - Z[k - 1]) * diff_z * 0.5;
        for (int i = 1; i < mpi_IMAX - 1; i++)
            for (int j = 1; j < mpi_JMAX - 1; j++)
                for (int k = 1; k < mpi_KMAX - 1; k++) {
                    double J = X_over_x[i] * Y_over_y[j] * Z_over_z[k] + Z_over_x * X_over_y * Y_over_z + Y_over_x * Z_over_y * X_over_z;
                    J += -X_over_x[i] * Z_over_y * Y_over_z - Y_over_x * X_over_y * Z_over_z[k] - Z_over_x * Y_over_y[j] * X_over_z;
                    Jacobian[i][j][k] = 1.0 / J;
                }
    }
    void load(double ***t, const std::vector<double> &f)
    {
        for (int k = 0; k < mpi_KMAX; ++k)
            for (int j = 0; j < mpi_JMAX; ++j)
                for (int i = 0; i < mpi_IMAX; ++i) t[i][j][k] = f[i + (size_t)mpi_IMAX * (j + (size_t)mpi_JMAX * k)];
    }
    std::vector<double> dump(double ***t)
    {
        std::vector<double> f((size_t)mpi_IMAX * mpi_JMAX * mpi_KMAX, 0.0);
        for (int k = 0; k < mpi_KMAX; ++k)
            for (int j = 0; j < mpi_JMAX; ++j)
                for (int i = 0; i < mpi_IMAX; ++i) f[i + (size_t)mpi_IMAX * (j + (size_t)mpi_JMAX * k)] = t[i][j][k];
        return f;
    }
    void fp() { porous::cal_fp(u_n, v_n, w_n, p_n); }
    void fuvw() { porous::cal_fuvw(u_n, v_n, w_n, p_n); }
    double ***un() { return u_n; } double ***vn() { return v_n; } double ***wn() { return w_n; } double ***pn() { return p_n; }
    double ***unn() { return u_nn; } double ***vnn() { return v_nn; } double ***wnn() { return w_nn; } double ***pnn() { return p_nn; }
};

template <class T> T *blank()
{
    void *mem = ::operator new(sizeof(T));
    memset(mem, 0, sizeof(T));
    return reinterpret_cast<T *>(mem);
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: ref_probe case.bin out.bin\n"); return 2; }
    MPI_Init(&argc, &argv);
    FILE *fin = fopen(argv[1], "rb");
    g_out = fopen(argv[2], "wb");
    if (!fin || !g_out) { fprintf(stderr, "cannot open files\n"); return 2; }
    Reader rd{fin};
    /* record 0: ints  nx ny nz hdr[10] bc[6] vmethod frameN counter pml_time jmf lead trail seed */
    std::vector<double> iv = rd.d();
    int nx = (int)iv[0], ny = (int)iv[1], nz = (int)iv[2];
    int hdr[10]; for (int q = 0; q < 10; ++q) hdr[q] = (int)iv[3 + q];
    int bc[6]; for (int q = 0; q < 6; ++q) bc[q] = (int)iv[13 + q];
    int vmethod = (int)iv[19], frameN = (int)iv[20], counter = (int)iv[21], pml_time = (int)iv[22];
    int jmf = (int)iv[23], lead = (int)iv[24], trail = (int)iv[25];
    unsigned long long seed = (unsigned long long)iv[26];
    /* record 1: doubles  dt dx dy dz gamma Pav c vcoef alpha[3] R[3] */
    std::vector<double> dv = rd.d();
    Grid g{nx, ny, nz, rd.d(), rd.d(), rd.d()};
    std::vector<double> fu = rd.d(), fv = rd.d(), fw = rd.d(), fp = rd.d();
    std::vector<double> fu2 = rd.d(), fv2 = rd.d(), fw2 = rd.d(), fp2 = rd.d();

    int dims[3] = {1, 1, 1}, per[3] = {0, 0, 0};
    MPI_Comm comm;
    MPI_Cart_create(MPI_COMM_WORLD, 3, dims, per, 1, &comm);

    AirProbe *a = blank<AirProbe>();
    a->setup(g, hdr, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], vmethod, dv[7], jmf, lead, trail, comm);
    /* PML judges from the absorbing faces (airpore.cpp:74-105, 182-247) */
    auto judge = [](bool hi, bool lo) { int n = (hi ? 1 : 0) + (lo ? 10 : 0); return n == 1 ? 2 : n == 10 ? 1 : n == 11 ? 3 : 0; };
    PML_boundary xb{judge(bc[4] == 2, bc[5] == 2), dv[8], dv[11], 0};
    PML_boundary yb{judge(bc[2] == 2, bc[3] == 2), dv[9], dv[12], pml_time};
    PML_boundary zb{judge(bc[0] == 2, bc[1] == 2), dv[10], dv[13], 0};
    a->metrics(0);
    a->pml(xb, yb, zb);
    a->qmove(0);
    a->fill_q(seed);
    a->emit_q("q0");
    if (frameN != 0) {
        a->set_gate(counter, (yb.judge == 1 || yb.judge == 3) ? hdr[4] : a->gateJ1());
        a->metrics(frameN);
        a->wind(frameN);
        a->qmove(frameN);
        a->emit_q("qshift");
    } else {
        a->wind(0);
    }
    emit("U", a->U()); emit("V", a->V()); emit("W", a->W());
    a->set_gate(counter, a->gateJ1());
    a->load(a->un(), fu); a->load(a->vn(), fv); a->load(a->wn(), fw); a->load(a->pn(), fp);
    a->fp();
    emit("dp", a->dump(a->pnn()));
    emit("gateJ1", std::vector<double>{(double)a->gateJ1()});
    a->fuvw();
    emit("du", a->dump(a->unn())); emit("dv", a->dump(a->vnn())); emit("dw", a->dump(a->wnn()));
    /* Q updates from fresh fields */
    a->load(a->pnn(), fp2);
    a->qp();
    a->emit_q("qp");
    a->load(a->unn(), fu2); a->load(a->vnn(), fv2); a->load(a->wnn(), fw2);
    a->qvw();
    a->emit_q("qvw");

    /* porous stencils on the same grid with the ground parameters of record 1 extras */
    std::vector<double> pp = rd.d(); /* q poro sigma */
    PoreProbe *po = blank<PoreProbe>();
    po->setup(g, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], pp[0], pp[1], pp[2]);
    po->load(po->un(), fu); po->load(po->vn(), fv); po->load(po->wn(), fw); po->load(po->pn(), fp);
    po->fp();
    emit("pore_dp", po->dump(po->pnn()));
    po->fuvw();
    emit("pore_du", po->dump(po->unn())); emit("pore_dv", po->dump(po->vnn())); emit("pore_dw", po->dump(po->wnn()));

    /* FFT of a series (record: m, then re, im) */
    std::vector<double> fm = rd.d();
    std::vector<double> re = rd.d(), im = rd.d();
    FFT(1, (long)fm[0], re.data(), im.data());
    emit("fft_re", re); emit("fft_im", im);
    std::vector<double> re2 = rd.d(), im2 = rd.d();
    FFT(-1, (long)fm[0], re2.data(), im2.data());
    emit("ifft_re", re2); emit("ifft_im", im2);
    if (argc > 3) FFT_output((long)fm[0], re.data(), im.data(), argv[3], dv[0]);

    /* Geo_over_var and mpe_decomp1d on a few arguments */
    std::vector<double> geo;
    for (int t = 0; t < 16; ++t) {
        double x[5] = {hashval(t * 5 + 1) * 1e3, hashval(t * 5 + 2) * 1e3, hashval(t * 5 + 3) * 1e3, hashval(t * 5 + 4) * 1e3, hashval(t * 5 + 5) * 1e3};
        geo.push_back(x[0]); geo.push_back(x[1]); geo.push_back(x[2]); geo.push_back(x[3]); geo.push_back(x[4]);
        geo.push_back(Geo_over_var(x[0], x[1], x[2], x[3], x[4]));
    }
    emit("geo", geo);
    std::vector<double> dec;
    for (int n = 1; n <= 40; n += 3)
        for (int np = 1; np <= 9; ++np)
            for (int id = 0; id < np; ++id) {
                int s, e;
                mpe_decomp1d(n, np, id, s, e);
                dec.push_back(n); dec.push_back(np); dec.push_back(id); dec.push_back(s); dec.push_back(e);
            }
    emit("decomp", dec);

    /* embedded obstacle (record: judge points coef[12] mat[3][3] box[6]); the material
     * parameters are derived as airpore.cpp:282-293 does, and the stencils rerun on a
     * fresh air object with the same PML auxiliaries as above */
    std::vector<double> ob = rd.d();
    if (ob[0] != 0) {
        AirProbe *b = blank<AirProbe>();
        b->setup(g, hdr, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], vmethod, dv[7], jmf, lead, trail, comm);
        b->metrics(0);
        b->pml(xb, yb, zb);
        b->qmove(0);
        b->fill_q(seed);
        b->wind(0);
        b->set_gate(counter, b->gateJ1());
        hillpore hp[3];
        memset(hp, 0, sizeof hp);
        double Aav_ = pow(dv[6], 2) / dv[4] / dv[5];
        for (int m = 0; m < 3; ++m) {
            hillpore &h = hp[m];
            h.curve_judge = (int)ob[0]; h.curve_points = (int)ob[1];
            double *cf[12] = {&h.curve_coefa, &h.curve_coefb, &h.curve_coefc, &h.curve_coefd, &h.curve_coefe, &h.curve_coeff,
                              &h.curve_coefg, &h.curve_coefh, &h.curve_coefi, &h.curve_coefj, &h.curve_coefk, &h.curve_coefl};
            for (int q = 0; q < 12; ++q) *cf[q] = ob[2 + q];
            h.q = ob[14 + 3 * m]; h.porosity = ob[15 + 3 * m]; h.resistivity = ob[16 + 3 * m];
            h.eff_density = h.q * h.q / h.porosity / Aav_;
            h.Kp = 1.0 / dv[4] / dv[5];
            h.curve_xstar = ob[23]; h.curve_xstop = ob[24]; h.curve_ystar = ob[25];
            h.curve_ystop = ob[26]; h.curve_zstar = ob[27]; h.curve_zstop = ob[28];
        }
        b->hill(hp[0], hp[1], hp[2]);
        emit("obox", b->hill_box());
        emit("omask", b->hill_mask());
        b->load(b->un(), fu); b->load(b->vn(), fv); b->load(b->wn(), fw); b->load(b->pn(), fp);
        b->fp();
        emit("odp", b->dump(b->pnn()));
        b->fuvw();
        emit("odu", b->dump(b->unn())); emit("odv", b->dump(b->vnn())); emit("odw", b->dump(b->wnn()));
    }
    fclose(g_out);
    fclose(fin);
    MPI_Finalize();
    return 0;
}
