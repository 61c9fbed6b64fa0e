// lee_model.h -- host-side FP64 model of one solver instance.
//
// Everything that is static per run or per moving frame is derived here, once, in FP64 on
// the host with the platform libm: validation of the deck, dimensions, PML slab geometry
// and per-index absorption profiles, porous and TDBC coefficients, wind fields, metrics,
// receiver indices and the frame schedule.  The device kernels then only add, subtract,
// multiply and divide, so the FP64 device path reproduces the reference arithmetic.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lee_abi.h"

namespace lee {

// One PML slab of auxiliary fields Q (air.h:56-58): `n` planes normal to `axis`,
// starting at global index `lo`.  Storage is x-fastest over the slab box.
struct SlabGeom {
    int axis = 0;     // 0 x, 1 y, 2 z
    int n = 0;        // planes (0 = inactive)
    int lo = 0;       // first global plane
    size_t count(int nx, int ny, int nz) const
    {
        size_t a = axis == 0 ? n : nx, b = axis == 1 ? n : ny, c = axis == 2 ? n : nz;
        return a * b * c;
    }
};

// Per-index absorption profile of one PML face: sigma, (1 - 1/beta), (1/beta - 1)
// (air.cpp:1580-1581 and the (1 -+ 1/beta) factors of air.cpp:1583-1905, 2048-2290).
struct Profile {
    std::vector<double> s, c, d;
};

// Terrain-following grid (extension): physical height Z(i,j,k) = Z[k] + h(i,j)*tk[k], with
// metric factors Z_over_z = Zz[k]*sh(i,j), Z_over_x = hx(i,j)*tk[k], Z_over_y = hy(i,j)*tk[k].
struct TerrainTables {
    std::vector<double> h, sh, hx, hy;    // [i + nx*j]
    std::vector<double> tk;               // [k]
};

struct FrameTables {
    std::vector<double> Y;                // air y coordinates of the frame (calculation_2D.cpp:410-415)
    std::vector<double> Xx, Yy, Zz;       // air metrics X_over_x ... (calculation_2D.cpp:417-419)
    std::vector<double> pXx, pYy, pZz;    // porous metrics
    Profile x1, x2, y1, y2, z1, z2;       // PML profiles for this frame
    std::vector<double> U, V, W;          // wind (empty when zero everywhere)
    TerrainTables ter, pter;              // air and porous terrain tables (empty when flat)
};

struct Model {
    lee_deck d{};
    lee_options opt{};
    std::vector<double> ax, ay, az;       // axisbound_X/Y/Z
    std::vector<double> pax, pay, paz;    // porous coordinates
    int nx = 0, ny = 0, nz = 0;           // air frame dims
    int pnx = 0, pny = 0, pnz = 0;        // porous dims
    int ijkmax = 0;
    double dxi = 0, dyi = 0, dzi = 0, dt = 0;
    double pdxi = 0, pdyi = 0, pdzi = 0;
    double gamma = 0, Pav = 0, Aav = 0, gP = 0;
    bool south_porous = false, east_tdbc = false, iface_copy = true, has_wind = false;
    // extensions: RK4, terrain-following grid, wind along y only (velocity_method 7)
    bool rk4 = false, terrain = false, wind_v_only = false;
    // embedded porous obstacle (air::cal_hill_var, air.cpp:938-1105): material id per node of
    // the box [ob[0], ob[1]] x [ob[2], ob[3]] x [ob[4], ob[5]] (global indices, i fastest);
    // id 1 = canopy, 2 = trunk (their coefficients at index 1, 2; index 0 is the hill's)
    bool obstacle = false;
    int ob[6] = {0, 0, 0, 0, 0, 0};
    std::vector<unsigned char> omask;
    double o_gama[3] = {}, o_invbeta[3] = {}, o_eff[3] = {};
    // PML (air.cpp:509-937)
    int judge[3] = {0, 0, 0};
    int I1 = 0, I2 = 0, J1 = 0, J2 = 0, J3 = 0, K1 = 0, K2 = 0;
    int pml_time = 0;
    double Rx = 0, Ry = 0, Rz = 0, alx = 0, aly = 0, alz = 0;
    double px1 = 0, px2 = 0, W1 = 0, W2 = 0, py1 = 0, py2 = 0, L1 = 0, L2 = 0, pz1 = 0, pz2 = 0, H1 = 0, H2 = 0;
    SlabGeom slab[6];                     // x1 x2 y1 y2 z1 z2
    // porous (porous.cpp:40-51)
    double p_eff = 0, p_gama = 0, p_invbeta = 0, p_beta = 0;
    // TDBC (calculation_2D.h:71-75, airpore.cpp:252-254)
    double Zinf = 0, tau = 0;
    // TDBC recursion constants: e^{-r_k dt/tau}, residues a_k, 1/(1 + sum_k a_k dt/tau)
    double td_decay[6] = {}, td_gain[6] = {}, td_inv = 0;
    // ground impedance on the South face (extension): the same recursion with K poles
    // closing the ground interface plane k = 0, which carries the porous equations of the
    // ground's top material (deck pore_q, pore_porosity, pore_resistivity)
    bool ground = false;
    int gK = 0;
    double gZinf = 0, g_decay[8] = {}, g_gain[8] = {}, g_inv = 0, g_tb = 0;
    double g_gama = 0, g_eff = 0, g_ib = 0;
    // moving frame
    int mf_input = 0, lead = 0, trail = 0;
    // receivers (whole-domain indices)
    std::vector<int> ri, rj, rk;
    std::vector<double> rpos;

    // Returns an empty string on success, else the reason; `code` receives LEE_OK,
    // LEE_ERR_UNSUPPORTED (a path the reference never implemented) or LEE_ERR_ARG.
    std::string build(const lee_deck &deck, const lee_options &o, int &code);
    // host_wind = false leaves the log wind of velocity_method 7 to the device (t.U/V/W empty)
    void frame_tables(int N, FrameTables &t, bool host_wind = true) const;
    std::string obstacle_mask();
    // terrain height of the hills at world (x, y)
    double hill_height(double x, double y) const;
    void terrain_tables(const std::vector<double> &X, const std::vector<double> &Y, const std::vector<double> &Z,
                        double dxi_, double dyi_, bool pore_layer, TerrainTables &t) const;
    void initial_condition(std::vector<double> &p, bool porous) const;
    int last_frame() const;
    int frame_count() const;
    int frame_steps(int mf, int temp_time) const;
    // linear index of receiver r in frame N (x-fastest), or -1 when it samples 0
    long long receiver_index(int r, int N) const;
};

}  // namespace lee
