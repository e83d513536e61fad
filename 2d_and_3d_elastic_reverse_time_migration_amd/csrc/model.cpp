// model.cpp - synthetic earth models and the reference's raw-float32 I/O.
//
// Files are raw little-endian float32 without a header, as the reference
// reads and writes them (3D_elastic_RTM/3D_zzzzz:3-139); 3D volumes are in
// disk order [nx][ny][nz] (z fastest).
//
// Synthetic models (seeded std::mt19937_64, so a seed fixes the model):
//  kind 0: three flat layers vp 2000/2800/3500 m/s at 1/3 and 2/3 depth plus
//          one Gaussian vp anomaly (+300 m/s, sigma 200 m) centred at half
//          width, 0.55 depth;
//  kind 1: five dipping layers with vp 1800, 2400, 3000, 3600, 4200 m/s whose
//          interfaces have a random common dip (0-10 degrees, random azimuth)
//          plus three random Gaussian vp anomalies (+-150..400 m/s,
//          sigma 100..300 m).
// Both use vs = vp/1.732 and density = 230*vp^0.25 (the reference's
// vp_set_density constant, 3D_elastic_RTM/zzzzz:1272-1287).
#include <cmath>
#include <cstdio>
#include <string>
#include <random>
#include <vector>

#include "common.hpp"

namespace erm {

int make_model3d(int kind, int nx, int ny, int nz, float dx, float dy, float dz, unsigned long long seed,
                 float* vp, float* vs, float* rho) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !vp) return ERM_E_PARAM;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double Lx = nx * (double)dx, Ly = ny * (double)dy, Lz = nz * (double)dz;
    struct Anom { double x, y, z, s, a; };
    std::vector<Anom> an;
    double dipx = 0, dipy = 0;
    std::vector<double> z0;
    std::vector<double> vl;
    if (kind == 0) {
        vl = {2000.0, 2800.0, 3500.0};
        z0 = {Lz / 3.0, 2.0 * Lz / 3.0};
        an.push_back({Lx / 2.0, Ly / 2.0, 0.55 * Lz, 200.0, 300.0});
    } else if (kind == 1) {
        vl = {1800.0, 2400.0, 3000.0, 3600.0, 4200.0};
        double dip = U(rng) * 10.0 * M_PI / 180.0, az = U(rng) * 2.0 * M_PI;
        dipx = std::tan(dip) * std::cos(az);
        dipy = std::tan(dip) * std::sin(az);
        for (int l = 1; l < 5; ++l) z0.push_back(Lz * (0.12 + 0.19 * l) + (U(rng) - 0.5) * 0.05 * Lz);
        for (int a = 0; a < 3; ++a) {
            double amp = 150.0 + 250.0 * U(rng);
            if (U(rng) < 0.5) amp = -amp;
            an.push_back({Lx * (0.15 + 0.7 * U(rng)), Ly * (0.15 + 0.7 * U(rng)), Lz * (0.25 + 0.6 * U(rng)),
                          100.0 + 200.0 * U(rng), amp});
        }
    } else {
        return ERM_E_PARAM;
    }
    const double xc = Lx / 2.0, yc = Ly / 2.0;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                double x = ix * (double)dx, y = iy * (double)dy, z = iz * (double)dz;
                double shift = dipx * (x - xc) + dipy * (y - yc);
                int l = 0;
                while (l < (int)z0.size() && z >= z0[l] + shift) ++l;
                double v = vl[l];
                for (const Anom& A : an) {
                    double r2 = (x - A.x) * (x - A.x) + (y - A.y) * (y - A.y) + (z - A.z) * (z - A.z);
                    v += A.a * std::exp(-r2 / (2.0 * A.s * A.s));
                }
                size_t n = ((size_t)ix * ny + iy) * nz + iz;
                vp[n] = (float)v;
                if (vs) vs[n] = (float)(v / 1.732);
                if (rho) rho[n] = (float)(1000.0 * 0.23 * std::sqrt(std::sqrt(v)));
            }
    return ERM_OK;
}

// Written to `path`.tmp and renamed into place, so a file under `path` is
// always complete (a run killed while writing leaves only the .tmp file).
int write_raw(const char* path, const float* data, size_t n) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return ERM_E_IO;
    size_t w = fwrite(data, sizeof(float), n, f);
    int rc = fclose(f);
    if (w != n || rc != 0 || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return ERM_E_IO;
    }
    return ERM_OK;
}

int read_raw(const char* path, float* data, size_t n) {
    FILE* f = fopen(path, "rb");
    if (!f) return ERM_E_IO;
    size_t r = fread(data, sizeof(float), n, f);
    fclose(f);
    return r == n ? ERM_OK : ERM_E_IO;
}

}  // namespace erm
