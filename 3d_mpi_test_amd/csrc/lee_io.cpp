// lee_io.cpp -- the drop-in file contract: input.txt + coordi_*.dat in, receiver series
// and spectra out, and the radix-2 FFT used for the spectra.
#include <sys/stat.h>
#include <sys/types.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lee_abi.h"

namespace lee {
void set_global_error(const std::string &m);
}

namespace {

int fail(int code, const std::string &msg)
{
    lee::set_global_error(msg);
    return code;
}

// Reads one coordi_<stem>{1,2,3}.dat triple (calculation_2D.cpp:44-68, 154-204).  `ny`
// values are taken from the second file (the frame length when the moving frame is on).
int read_grid(const std::string &dir, const char *stem, int hdr[10], int mf_judge, int mf_jmax,
              double **X, double **Y, double **Z)
{
    std::string f1 = dir + "/" + stem + "1.dat", f2 = dir + "/" + stem + "2.dat", f3 = dir + "/" + stem + "3.dat";
    std::ifstream a(f1.c_str(), std::ios::in | std::ios::binary);
    if (!a) return fail(LEE_ERR_IO, "cannot open file: " + f1);
    for (int q = 0; q < 10; ++q)
        if (!(a >> hdr[q])) return fail(LEE_ERR_IO, "malformed header in " + f1);
    int nx = hdr[0], nyf = hdr[3], nz = hdr[6];
    int ny = mf_judge == 1 ? mf_jmax : nyf;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > (1 << 20) || ny > (1 << 20) || nz > (1 << 20))
        return fail(LEE_ERR_IO, "implausible grid size in " + f1);
    double *x = (double *)std::malloc(sizeof(double) * nx);
    double *y = (double *)std::malloc(sizeof(double) * ny);
    double *z = (double *)std::malloc(sizeof(double) * nz);
    if (!x || !y || !z) { std::free(x); std::free(y); std::free(z); return fail(LEE_ERR_IO, "out of memory"); }
    auto bad = [&](const std::string &m) { std::free(x); std::free(y); std::free(z); return fail(LEE_ERR_IO, m); };
    for (int i = 0; i < nx; ++i) if (!(a >> x[i])) return bad("too few X values in " + f1);
    std::ifstream b(f2.c_str(), std::ios::in | std::ios::binary);
    if (!b) return bad("cannot open file: " + f2);
    for (int j = 0; j < ny; ++j) if (!(b >> y[j])) return bad("too few Y values in " + f2);
    std::ifstream c(f3.c_str(), std::ios::in | std::ios::binary);
    if (!c) return bad("cannot open file: " + f3);
    for (int k = 0; k < nz; ++k) if (!(c >> z[k])) return bad("too few Z values in " + f3);
    *X = x; *Y = y; *Z = z;
    return LEE_OK;
}

bool make_dir(const std::string &p)
{
    struct stat st;
    if (stat(p.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
    return mkdir(p.c_str(), 0777) == 0;
}

std::string series_name(const std::string &dir, const char *stem, int i)
{
    char b[64];
    std::snprintf(b, sizeof b, "%s%04d.dat", stem, i);
    return dir + "/solution/" + b;
}

// Radix-2 decimation-in-time FFT, in place, N = 2^m: bit-reversal permutation followed by
// log2(N) butterfly passes whose twiddle factor is advanced by complex multiplication with
// the half-angle root (cos, sin) obtained from the previous one by the half-angle formulas.
// dir = 1 forward (exp(-i w t), result scaled by 1/N), dir = -1 inverse.  Same arithmetic
// as FFT() of FFT.cpp:15-82, so spectra agree bit for bit with the reference's.
void fft_inplace(int dir, int m, double *re, double *im)
{
    long n = 1L << m;
    long half = n >> 1, j = 0;
    for (long i = 0; i < n - 1; ++i) {
        if (i < j) { std::swap(re[i], re[j]); std::swap(im[i], im[j]); }
        long k = half;
        while (k <= j) { j -= k; k >>= 1; }
        j += k;
    }
    double cr = -1.0, ci = 0.0;  // root of unity of the current pass
    for (long span = 1; span < n; span <<= 1) {
        double wr = 1.0, wi = 0.0;
        for (long o = 0; o < span; ++o) {
            for (long a = o; a < n; a += 2 * span) {
                long b = a + span;
                double tr = wr * re[b] - wi * im[b];
                double ti = wr * im[b] + wi * re[b];
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr; im[a] += ti;
            }
            double t = wr * cr - wi * ci;
            wi = wr * ci + wi * cr;
            wr = t;
        }
        ci = std::sqrt((1.0 - cr) / 2.0);
        if (dir == 1) ci = -ci;
        cr = std::sqrt((1.0 + cr) / 2.0);
    }
    if (dir == 1)
        for (long i = 0; i < n; ++i) { re[i] /= n; im[i] /= n; }
}

}  // namespace

extern "C" {

// airpore::airpore (airpore.cpp:17-316): positional values read with operator>> (which
// skips blank lines), the rest of each line dropped with ignore(100,'\n').
int lee_deck_load(const char *dir, lee_deck *out)
{
    if (!dir || !out) return fail(LEE_ERR_ARG, "NULL argument");
    std::memset(out, 0, sizeof *out);
    std::string d(dir);
    std::string path = d + "/input.txt";
    std::ifstream in(path.c_str());
    if (!in) return fail(LEE_ERR_IO, "cannot open " + path);
    lee_deck &k = *out;
    auto skip = [&]() { in.ignore(100, '\n'); };
    in >> k.gauss_width; skip();
    in >> k.mf_jmax >> k.mf_lead_dj >> k.mf_trail_dj >> k.mf_judge >> k.mf_ratio; skip();
    in >> k.scheme; skip();
    for (int i = 0; i < 6; ++i) in >> k.bc[i];
    skip(); skip();
    in >> k.time_domain >> k.out_difft >> k.fft_m >> k.frequency; skip();
    in >> k.velocity_coef >> k.velocity_method; skip();
    in >> k.gamma >> k.Pav >> k.c; skip();
    in >> k.pml_alpha[0] >> k.pml_R[0]; skip();
    in >> k.pml_alpha[1] >> k.pml_R[1] >> k.pml_y_time; skip();
    in >> k.pml_alpha[2] >> k.pml_R[2]; skip();
    in >> k.pore_q >> k.pore_porosity >> k.pore_resistivity; skip();
    in >> k.src[0] >> k.src[1] >> k.src[2] >> k.rcv[0] >> k.rcv[1] >> k.rcv[2]; skip();
    in >> k.dt; skip();
    in >> k.dx >> k.dy >> k.dz; skip();
    in >> k.mpi_dims[0] >> k.mpi_dims[1] >> k.mpi_dims[2]; skip();
    skip(); skip();
    std::string restart_file;
    in >> k.restart >> k.restart_out >> restart_file;
    skip(); skip();
    // embedded porous obstacle (airpore.cpp:270-299): curve line, hill / canopy / trunk
    // materials, bounding box
    in >> k.hill_judge;
    for (int q = 0; q < 12; ++q) in >> k.hill_coef[q];
    in >> k.hill_points; skip();
    for (int q = 0; q < 3; ++q) { in >> k.hill_mat[q][0] >> k.hill_mat[q][1] >> k.hill_mat[q][2]; skip(); }
    for (int q = 0; q < 6; ++q) in >> k.hill_box[q];
    skip();
    skip(); skip();
    if (k.bc[LEE_FACE_S] == LEE_BC_POROUS) {
        in >> k.pore_dx >> k.pore_dy >> k.pore_dz; skip();
    }
    if (in.fail()) return fail(LEE_ERR_IO, "malformed input.txt (a positional value failed to parse)");
    double *x = nullptr, *y = nullptr, *z = nullptr;
    int rc = read_grid(d, "coordi_air", k.air_hdr, k.mf_judge, k.mf_jmax, &x, &y, &z);
    if (rc) return rc;
    k.air_x = x; k.air_y = y; k.air_z = z;
    if (k.bc[LEE_FACE_S] == LEE_BC_POROUS) {
        rc = read_grid(d, "coordi_pore", k.pore_hdr, k.mf_judge, k.mf_jmax, &x, &y, &z);
        if (rc) { lee_deck_free(out); return rc; }
        k.pore_x = x; k.pore_y = y; k.pore_z = z;
        k.has_pore = 1;
    }
    return LEE_OK;
}

void lee_deck_free(lee_deck *deck)
{
    if (!deck) return;
    std::free((void *)deck->air_x); std::free((void *)deck->air_y); std::free((void *)deck->air_z);
    std::free((void *)deck->pore_x); std::free((void *)deck->pore_y); std::free((void *)deck->pore_z);
    deck->air_x = deck->air_y = deck->air_z = nullptr;
    deck->pore_x = deck->pore_y = deck->pore_z = nullptr;
    deck->has_pore = 0;
}

void lee_options_default(lee_options *opt)
{
    if (!opt) return;
    std::memset(opt, 0, sizeof *opt);
    opt->precision = 64;
}

int lee_fft(int dir, int m, double *re, double *im)
{
    if (!re || !im || m < 0 || m > 30 || (dir != 1 && dir != -1)) return fail(LEE_ERR_ARG, "bad FFT arguments");
    fft_inplace(dir, m, re, im);
    return LEE_OK;
}

// ./solution/pty%04d.dat, one value per line in ios::scientific with precision 6
// (airpore.cpp:468-531, 577).
int lee_write_receivers(const char *dir, const double *rec, int nsteps, int nrec)
{
    if (!dir || (!rec && nsteps > 0) || nsteps < 0 || nrec < 0) return fail(LEE_ERR_ARG, "bad arguments");
    std::string base(dir);
    if (!make_dir(base + "/solution")) return fail(LEE_ERR_IO, "cannot create " + base + "/solution");
    for (int r = 0; r < nrec; ++r) {
        std::ofstream o(series_name(base, "pty", r).c_str(), std::ios::out | std::ios::trunc);
        if (!o) return fail(LEE_ERR_IO, "cannot write " + series_name(base, "pty", r));
        o.setf(std::ios::scientific, std::ios::floatfield);
        o.precision(6);
        for (int t = 0; t < nsteps; ++t) o << rec[(size_t)t * nrec + r] << std::endl;
    }
    return LEE_OK;
}

// airpore::get_FFT_y1 (airpore.cpp:966-1011) + FFT_output (FFT.cpp:84-121): each series is
// taken as written to its pty file (6-digit text), zero-padded to 2^fft_m, transformed,
// and written as "f |X(f)|" at setprecision(10), N/2+1 lines.  The reference always
// processes five files; missing series read as zeros.
int lee_write_spectra(const char *dir, const double *rec, int nsteps, int nrec, int fft_m, double dt)
{
    if (!dir || (!rec && nsteps > 0) || nsteps < 0 || nrec < 0 || fft_m < 1 || fft_m > 26 || !(dt > 0))
        return fail(LEE_ERR_ARG, "bad arguments");
    long n = 1L << fft_m;
    if (n < nsteps) return fail(LEE_ERR_ARG, "2^FFT_m is shorter than the series (the reference overflows)");
    std::string base(dir);
    if (!make_dir(base + "/solution")) return fail(LEE_ERR_IO, "cannot create " + base + "/solution");
    int nfiles = nrec < 5 ? 5 : nrec;
    std::vector<double> re(n), im(n);
    for (int r = 0; r < nfiles; ++r) {
        std::fill(re.begin(), re.end(), 0.0);
        std::fill(im.begin(), im.end(), 0.0);
        if (r < nrec) {
            for (int t = 0; t < nsteps; ++t) {
                std::ostringstream s;
                s.setf(std::ios::scientific, std::ios::floatfield);
                s.precision(6);
                s << rec[(size_t)t * nrec + r];
                re[t] = std::strtod(s.str().c_str(), nullptr);
            }
        }
        fft_inplace(1, fft_m, re.data(), im.data());
        std::ofstream o(series_name(base, "p_f2_", r).c_str(), std::ios::out | std::ios::trunc | std::ios::binary);
        if (!o) return fail(LEE_ERR_IO, "cannot write " + series_name(base, "p_f2_", r));
        o << std::setprecision(10) << "0.0 " << std::abs(std::complex<double>(re[0], im[0])) << std::endl;
        long jj;
        for (jj = 1; jj <= n / 2 - 1; ++jj)
            o << std::setprecision(10) << (jj / dt / n) << " " << std::abs(std::complex<double>(re[jj], im[jj])) << std::endl;
        o << std::setprecision(10) << (jj / dt / n) << " " << std::abs(std::complex<double>(re[n / 2], im[n / 2])) << std::endl;
    }
    return LEE_OK;
}

}  // extern "C"
