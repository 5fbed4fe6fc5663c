// rgb_uplift.cpp -- Jakob-Hanika RGB-to-spectrum table: generation, lookup and
// the generator's ".coeff" file format (host code).
//
// The table maps a Rec.709 colour to three coefficients of the sigmoid
// polynomial spectrum S(lambda) = sigmoid(c0 t^2 + c1 t + c2), t = (lambda -
// 380) / 400.  Its cells are the solutions of a small nonlinear least-squares
// problem (integrate S against the CIE 1931 observer under a flat
// illuminant, convert XYZ to Rec.709, match the cell's colour), solved by
// Gauss-Newton with a finite-difference Jacobian, Levenberg-Marquardt damping
// and a halving line search, stopping below 1e-4 (Jakob & Hanika 2019, §3;
// the reference's tools/JakobHanikaLUTGen.cpp:258-420).  Because the solver
// stops at that tolerance and warm-starts each cell from its neighbour, the
// table depends on the exact steps, seeds and walk order, which are therefore
// those of the reference's generator (:560-640), evaluated in the same
// floating-point order (the library is built with -ffp-contract=off).
#include "rise_hip_rgb_uplift.h"

#include "api_common.h"
#include "cie_table.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace rise_hip;

struct rise_hip_rgb_lut {
    int res = 0;
    std::vector<float> coeffs;  // 3 per cell
};

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kNLambda = 81;  // 380..780 nm at 5 nm
const double kCieX[kNLambda] = {RISE_CIE_X_VALUES};
const double kCieY[kNLambda] = {RISE_CIE_Y_VALUES};
const double kCieZ[kNLambda] = {RISE_CIE_Z_VALUES};

// XYZ (D65) -> linear Rec.709 (the reference's mxXYZtoRec709, Color.cpp).
const double kXyzToRec709[3][3] = {
    {3.240479, -1.537150, -0.498535}, {-0.969256, 1.875992, 0.041556}, {0.055648, -0.204043, 1.057311}};

constexpr double kAcceptTol = 1e-4;
constexpr int kMaxIter = 200;
constexpr double kFdStep = 5e-4;

inline double sigmoid(double x) { return 0.5 + x / (2.0 * std::sqrt(1.0 + x * x)); }

inline double spectrum(const double c[3], double lambda)
{
    const double t = (lambda - 380.0) * (1.0 / 400.0);
    return sigmoid(c[0] * t * t + c[1] * t + c[2]);
}

// The colour of spectrum c under a flat illuminant, normalized so that S = 1
// has Y = 1, in linear Rec.709.
void spectrum_to_rgb(const double c[3], double rgb[3])
{
    double X = 0.0, Y = 0.0, Z = 0.0, ny = 0.0;
    for (int i = 0; i < kNLambda; ++i) {
        const double s = spectrum(c, 380.0 + i * 5.0);
        X += s * kCieX[i];
        Y += s * kCieY[i];
        Z += s * kCieZ[i];
        ny += kCieY[i];
    }
    const double inv = 1.0 / ny;
    X *= inv;
    Y *= inv;
    Z *= inv;
    for (int k = 0; k < 3; ++k) rgb[k] = kXyzToRec709[k][0] * X + kXyzToRec709[k][1] * Y + kXyzToRec709[k][2] * Z;
}

double residual(const double c[3], const double target[3], double r[3])
{
    double rgb[3];
    spectrum_to_rgb(c, rgb);
    for (int k = 0; k < 3; ++k) r[k] = rgb[k] - target[k];
    return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

// Minimizes |rgb(c) - target| from the seed in c; leaves the best point found
// in c and returns its residual norm.
double solve(const double target[3], double c[3])
{
    double best[3] = {c[0], c[1], c[2]};
    double r[3];
    double best_norm = residual(c, target, r);
    for (int iter = 0; iter < kMaxIter; ++iter) {
        const double norm = residual(c, target, r);
        if (norm < best_norm) {
            best_norm = norm;
            std::copy(c, c + 3, best);
        }
        if (norm < kAcceptTol) break;
        // Central-difference Jacobian J[i][j] = d r_i / d c_j.
        double J[3][3];
        for (int j = 0; j < 3; ++j) {
            double cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
            cp[j] += kFdStep;
            cm[j] -= kFdStep;
            double fp[3], fm[3];
            spectrum_to_rgb(cp, fp);
            spectrum_to_rgb(cm, fm);
            for (int i = 0; i < 3; ++i) J[i][j] = (fp[i] - fm[i]) / (2.0 * kFdStep);
        }
        // Damped normal equations (J^T J + mu I) delta = J^T r.
        double A[3][3], g[3];
        for (int i = 0; i < 3; ++i) {
            g[i] = J[0][i] * r[0] + J[1][i] * r[1] + J[2][i] * r[2];
            for (int j = 0; j < 3; ++j) A[i][j] = J[0][i] * J[0][j] + J[1][i] * J[1][j] + J[2][i] * J[2][j];
        }
        const double mu = 1e-6 * std::max({A[0][0], A[1][1], A[2][2]});
        for (int i = 0; i < 3; ++i) A[i][i] += mu;
        const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                           A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        if (std::fabs(det) < 1e-40) break;
        const double id = 1.0 / det;
        // Inverse by the adjugate.
        const double inv[3][3] = {
            {(A[1][1] * A[2][2] - A[1][2] * A[2][1]) * id, -(A[0][1] * A[2][2] - A[0][2] * A[2][1]) * id,
             (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id},
            {-(A[1][0] * A[2][2] - A[1][2] * A[2][0]) * id, (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id,
             -(A[0][0] * A[1][2] - A[0][2] * A[1][0]) * id},
            {(A[1][0] * A[2][1] - A[1][1] * A[2][0]) * id, -(A[0][0] * A[2][1] - A[0][1] * A[2][0]) * id,
             (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id}};
        double delta[3];
        for (int i = 0; i < 3; ++i) delta[i] = inv[i][0] * g[0] + inv[i][1] * g[1] + inv[i][2] * g[2];
        // Halving line search: the first of 1, 1/2, ..., 1/512 that lowers the residual.
        bool moved = false;
        double step = 1.0;
        for (int ls = 0; ls < 10 && !moved; ++ls, step *= 0.5) {
            const double trial[3] = {c[0] - step * delta[0], c[1] - step * delta[1], c[2] - step * delta[2]};
            double rt[3];
            const double tn = residual(trial, target, rt);
            if (tn < norm) {
                moved = true;
                std::copy(trial, trial + 3, c);
                if (tn < best_norm) {
                    best_norm = tn;
                    std::copy(c, c + 3, best);
                }
            }
        }
        if (!moved) break;
    }
    std::copy(best, best + 3, c);
    return best_norm;
}

// z(i) = sin^2(pi/2 * i / (res - 1)): finer steps toward bright colours.
double grid_z(int i, int res)
{
    const double s = std::sin(0.5 * kPi * (double(i) / double(res - 1)));
    return s * s;
}

struct PlaneStats {
    int64_t converged = 0, failed = 0;
    double max_res = 0.0, sum_res = 0.0;
};

// One (max channel, z) plane: cells walked x-major, each seeded from the
// previous cell, the first of every row from the first of the previous row;
// a cell left above the tolerance is retried from (0, 0, 0).
void solve_plane(int maxc, int iz, int res, float* out, PlaneStats& st)
{
    const double z = grid_z(iz, res);
    double row_seed[3] = {0.0, 0.0, 0.0};
    for (int ix = 0; ix < res; ++ix) {
        const double x = double(ix) / double(res - 1);
        double prev[3] = {row_seed[0], row_seed[1], row_seed[2]};
        for (int iy = 0; iy < res; ++iy) {
            const double y = double(iy) / double(res - 1);
            double rgb[3];
            rgb[maxc] = z;
            rgb[(maxc + 1) % 3] = x * z;
            rgb[(maxc + 2) % 3] = y * z;
            double c[3] = {prev[0], prev[1], prev[2]};
            double rn;
            if (z < 1e-9) {
                // Black: no sigmoid reaches 0; a far-negative constant term
                // gives the nearest spectrum.
                c[0] = c[1] = 0.0;
                c[2] = -100.0;
                rn = 0.0;
            } else {
                rn = solve(rgb, c);
                if (rn > kAcceptTol) {
                    double cold[3] = {0.0, 0.0, 0.0};
                    const double rc = solve(rgb, cold);
                    if (rc < rn) {
                        std::copy(cold, cold + 3, c);
                        rn = rc;
                    }
                }
            }
            if (rn < kAcceptTol) ++st.converged;
            else ++st.failed;
            st.max_res = std::max(st.max_res, rn);
            st.sum_res += rn;
            float* cell = out + 3 * ((((size_t)maxc * res + iz) * res + ix) * res + iy);
            cell[0] = (float)c[0];
            cell[1] = (float)c[1];
            cell[2] = (float)c[2];
            std::copy(c, c + 3, prev);
            if (iy == 0) std::copy(c, c + 3, row_seed);
        }
    }
}

const float* cell_of(const rise_hip_rgb_lut& L, int maxc, int iz, int ix, int iy)
{
    const int r = L.res;
    iz = std::max(0, std::min(r - 1, iz));
    ix = std::max(0, std::min(r - 1, ix));
    iy = std::max(0, std::min(r - 1, iy));
    return L.coeffs.data() + 3 * ((((size_t)maxc * r + iz) * r + ix) * r + iy);
}

inline float lerp1(float a, float b, double t) { return float(double(a) * (1.0 - t) + double(b) * t); }

}  // namespace

extern "C" rise_hip_status rise_hip_rgb_lut_generate(int32_t resolution, int32_t threads, rise_hip_rgb_lut** out,
                                                     rise_hip_rgb_lut_stats* stats)
{
    clear_error();
    if (!out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "out_lut is NULL");
    *out = nullptr;
    if (resolution < 8 || resolution > 256) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "resolution outside [8, 256]");
    const auto t0 = std::chrono::steady_clock::now();
    rise_hip_rgb_lut* L = new (std::nothrow) rise_hip_rgb_lut;
    if (!L) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    try {
        L->res = resolution;
        L->coeffs.assign((size_t)3 * 3 * resolution * resolution * resolution, 0.0f);
        const int planes = 3 * resolution;
        std::vector<PlaneStats> ps((size_t)planes);
        std::atomic<int> next(0);
        auto worker = [&]() {
            for (int p; (p = next.fetch_add(1)) < planes;) solve_plane(p / resolution, p % resolution, resolution,
                                                                       L->coeffs.data(), ps[(size_t)p]);
        };
        int nt = threads > 0 ? threads : host_threads();
        nt = std::max(1, std::min(nt, planes));
        std::vector<std::thread> th;
        for (int k = 1; k < nt; ++k) th.emplace_back(worker);
        worker();
        for (auto& t : th) t.join();
        if (stats) {
            *stats = rise_hip_rgb_lut_stats{};
            stats->cells = (int64_t)3 * resolution * resolution * resolution;
            double sum = 0.0;
            for (const PlaneStats& s : ps) {
                stats->converged += s.converged;
                stats->failed += s.failed;
                stats->max_residual = std::max(stats->max_residual, s.max_res);
                sum += s.sum_res;
            }
            stats->mean_residual = sum / (double)stats->cells;
            stats->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    } catch (const std::bad_alloc&) {
        delete L;
        return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::system_error&) {
        delete L;
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "could not start threads");
    }
    *out = L;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_rgb_lut_read(const char* path, rise_hip_rgb_lut** out)
{
    clear_error();
    if (!path || !out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
    char magic[4];
    uint32_t h[4];
    if (!f.read(magic, 4) || std::memcmp(magic, "RJHL", 4) != 0 || !f.read((char*)h, sizeof(h)) ||
        h[0] != 0x00010000u || h[1] < 8 || h[1] > 256 || h[2] != 3 || h[3] != 3)
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": not an RJHL version-1 table");
    rise_hip_rgb_lut* L = new (std::nothrow) rise_hip_rgb_lut;
    if (!L) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "host allocation failed");
    L->res = (int)h[1];
    L->coeffs.resize((size_t)9 * L->res * L->res * L->res);
    if (!f.read((char*)L->coeffs.data(), (std::streamsize)(L->coeffs.size() * sizeof(float)))) {
        delete L;
        return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string(path) + ": table body truncated");
    }
    *out = L;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_rgb_lut_write(const rise_hip_rgb_lut* L, const char* path)
{
    clear_error();
    if (!L || !path) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot create ") + path);
    const uint32_t h[4] = {0x00010000u, (uint32_t)L->res, 3u, 3u};
    bool ok = std::fwrite("RJHL", 1, 4, f) == 4 && std::fwrite(h, 4, 4, f) == 4 &&
              std::fwrite(L->coeffs.data(), sizeof(float), L->coeffs.size(), f) == L->coeffs.size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RISE_HIP_OK : fail(RISE_HIP_ERR_INVALID_ARGUMENT, std::string("cannot write ") + path);
}

extern "C" rise_hip_status rise_hip_rgb_lut_get(const rise_hip_rgb_lut* L, int32_t* resolution, const float** coeffs)
{
    clear_error();
    if (!L) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL lut");
    if (resolution) *resolution = L->res;
    if (coeffs) *coeffs = L->coeffs.data();
    return RISE_HIP_OK;
}

// RGBToSpectrumTable::operator() (RGBToSpectrumTable.cpp:168-243).
extern "C" rise_hip_status rise_hip_rgb_lut_lookup(const rise_hip_rgb_lut* L, const double rgb_in[3],
                                                   float coeffs[3])
{
    clear_error();
    if (!L || !rgb_in || !coeffs) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    double v[3];
    for (int k = 0; k < 3; ++k) v[k] = std::max(0.0, std::min(1.0, rgb_in[k]));
    const double r = v[0], g = v[1], b = v[2];
    // The largest channel picks the sub-table; the next two channels in
    // cyclic order, divided by it, are the in-plane coordinates.
    int maxc = (r >= g && r >= b) ? 0 : (g >= r && g >= b) ? 1 : 2;
    const double z = v[maxc];
    const double x = z > 1e-9 ? v[(maxc + 1) % 3] / z : 0.0;
    const double y = z > 1e-9 ? v[(maxc + 2) % 3] / z : 0.0;
    const int res = L->res;
    const double top = (double)(res - 1);
    // Inverse of the sin^2 z grid.
    const double fz_idx = std::max(0.0, std::min(top, std::asin(std::sqrt(std::max(0.0, std::min(1.0, z)))) *
                                                          (2.0 / kPi) * top));
    const double fx_idx = std::max(0.0, std::min(top, x * top));
    const double fy_idx = std::max(0.0, std::min(top, y * top));
    const int iz = (int)std::floor(fz_idx), ix = (int)std::floor(fx_idx), iy = (int)std::floor(fy_idx);
    const double fz = fz_idx - iz, fx = fx_idx - ix, fy = fy_idx - iy;
    for (int k = 0; k < 3; ++k) {
        // Trilinear: along x, then y, then z.
        const float a0 = lerp1(cell_of(*L, maxc, iz, ix, iy)[k], cell_of(*L, maxc, iz, ix + 1, iy)[k], fx);
        const float a1 = lerp1(cell_of(*L, maxc, iz + 1, ix, iy)[k], cell_of(*L, maxc, iz + 1, ix + 1, iy)[k], fx);
        const float b0 = lerp1(cell_of(*L, maxc, iz, ix, iy + 1)[k], cell_of(*L, maxc, iz, ix + 1, iy + 1)[k], fx);
        const float b1 =
            lerp1(cell_of(*L, maxc, iz + 1, ix, iy + 1)[k], cell_of(*L, maxc, iz + 1, ix + 1, iy + 1)[k], fx);
        coeffs[k] = lerp1(lerp1(a0, b0, fy), lerp1(a1, b1, fy), fz);
    }
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_rgb_lut_release(rise_hip_rgb_lut* L)
{
    clear_error();
    delete L;
    return RISE_HIP_OK;
}

extern "C" rise_hip_status rise_hip_rgb_lut_default(const rise_hip_rgb_lut** out)
{
    clear_error();
    if (!out) return fail(RISE_HIP_ERR_INVALID_ARGUMENT, "NULL argument");
    static std::once_flag once;
    static rise_hip_rgb_lut* table = nullptr;
    static std::string why;
    std::call_once(once, [] {
        if (rise_hip_rgb_lut_generate(64, 0, &table, nullptr) != RISE_HIP_OK) why = rise_hip_last_error();
    });
    if (!table) return fail(RISE_HIP_ERR_OUT_OF_MEMORY, "RGB-to-spectrum table generation failed: " + why);
    *out = table;
    return RISE_HIP_OK;
}

// RGBSigmoidPolynomial::Eval (RGBSigmoidPolynomial.h:55-69) times scale.
extern "C" double rise_hip_sigmoid_spectrum_eval(const float c[3], double scale, double lambda_nm)
{
    if (!(lambda_nm >= 380.0 && lambda_nm <= 780.0)) return 0.0;
    const double t = (lambda_nm - 380.0) / 400.0;
    const double x = (double)c[0] * t * t + (double)c[1] * t + (double)c[2];
    return scale * (0.5 + x / (2.0 * std::sqrt(1.0 + x * x)));
}
