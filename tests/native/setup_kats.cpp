// Known-answer tests of the engine's float64 set-up functions (host code, no GPU), in the manner of
// the reference's own C++ tests: the bicubic upsample sample, the exposure factor, the paper
// curve model, the scan spectra, the dichroic filters, the Young-van Vliet coefficients and the
// geometry sizes (reference tests/test_print.cpp:202-258, tests/test_fx.cpp:140-149,
// tests/test_film.cpp:282-315).  Built and run by tests/test_setup_native.py:
//     setup_kats <resources folder> <paper stock id>
// prints one line per check and exits non-zero if any fails.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "setup.h"

using namespace spk;
using namespace spk::setup;

static int g_failed = 0;
#define CHECK(cond)                                                            \
    do {                                                                       \
        const bool ok_ = (cond);                                               \
        std::printf("%s %s:%d %s\n", ok_ ? "ok  " : "FAIL", __FILE__, __LINE__, #cond); \
        if (!ok_) ++g_failed;                                                  \
    } while (0)
#define NEAR(a, b, tol) CHECK(std::fabs(double(a) - double(b)) <= (tol))

static void mitchell_kats() {
    // a constant table is reproduced exactly; at a node a linear ramp stays a ramp; the end samples
    // mirror (weights 1/18, 16/18, 1/18 of rows 1, 0, 1 at t = 0 and columns 6, 7, 6 at t = 1)
    const uint32_t L = 8;
    std::vector<double> tab(L * L * 3);
    for (uint32_t i = 0; i < L; ++i)
        for (uint32_t j = 0; j < L; ++j) {
            tab[(i * L + j) * 3 + 0] = 0.25;
            tab[(i * L + j) * 3 + 1] = double(i);
            tab[(i * L + j) * 3 + 2] = 2.0 * double(j);
        }
    const Vec3 v = mitchell_sample(tab, L, 3.0 / 7.0, 4.0 / 7.0);
    NEAR(v[0], 0.25, 1e-15);
    NEAR(v[1], 3.0, 1e-14);
    NEAR(v[2], 8.0, 1e-14);
    const Vec3 e = mitchell_sample(tab, L, 0.0, 1.0);
    NEAR(e[1], 2.0 / 18.0, 1e-14);
    NEAR(e[2], (12.0 + 16.0 * 14.0 + 12.0) / 18.0, 1e-13);
}

static void exposure_kats() {
    // the reciprocal geometric mean of R_c = sum_l 10^-D(l) I(l) S_c(l): with D = 1, I = 1 and
    // S = 1, 2, 4 over 81 bands, R = 8.1, 16.2, 32.4 and the mean is 16.2; a NaN band drops out
    Spectra3 sens{};
    Spectrum light{}, dens{};
    for (size_t l = 0; l < 81; ++l) {
        light[l] = 1.0;
        dens[l] = 1.0;
        sens[3 * l] = 1.0;
        sens[3 * l + 1] = 2.0;
        sens[3 * l + 2] = 4.0;
    }
    NEAR(exposure_gain(dens, light, sens), 1.0 / (81.0 * 0.1 * 2.0), 1e-14);
    dens[10] = std::nan("");
    NEAR(exposure_gain(dens, light, sens), 1.0 / (80.0 * 0.1 * 2.0), 1e-14);
}

static void filter_kats() {
    const std::array<Spectrum, 3> f = dichroics();
    bool in01 = true;
    for (const Spectrum& s : f)
        for (double v : s) in01 = in01 && v >= 0.0 && v <= 1.0;
    CHECK(in01);
    // cyan passes blue and cuts red; yellow cuts blue and passes red; magenta cuts the middle
    CHECK(f[0][4] > 0.99 && f[0][76] < 0.01);     // 400 nm, 760 nm
    CHECK(f[2][4] < 0.01 && f[2][76] > 0.99);
    CHECK(f[1][34] < 0.01 && f[1][4] > 0.99 && f[1][76] > 0.99);  // 550 nm
    Spectrum lamp{};
    for (size_t l = 0; l < 81; ++l) lamp[l] = 0.5 + 0.01 * double(l);
    const Spectrum open = through_filters(lamp, f, {0.0, 0.0, 0.0});
    bool same = true;
    for (size_t l = 0; l < 81; ++l) same = same && open[l] == lamp[l];
    CHECK(same);  // 0 CC: tau = 1, every factor exactly 1
    const Spectrum y100 = through_filters(lamp, f, {0.0, 0.0, 100.0});  // 100 CC: tau = 0.1
    NEAR(y100[4], lamp[4] * (1.0 - (1.0 - f[2][4]) * 0.9), 1e-15);
}

static void yvv_kats() {
    for (double s : {0.8, 2.0, 2.6, 8.0, 32.0}) {
        const double q = s >= 2.5 ? 0.98711 * s - 0.96330 : 3.97156 - 4.14554 * std::sqrt(1.0 - 0.26891 * s);
        const double b0 = 1.57825 + 2.44413 * q + 1.4281 * q * q + 0.422205 * q * q * q;
        const double b1 = 2.44413 * q + 2.85619 * q * q + 1.26661 * q * q * q;
        const double b2 = -(1.4281 * q * q + 1.26661 * q * q * q);
        const double b3 = 0.422205 * q * q * q;
        const YvvCoeffs c = yvv(s);
        NEAR(c.b1, b1 / b0, 1e-12);
        NEAR(c.b2, b2 / b0, 1e-12);
        NEAR(c.b3, b3 / b0, 1e-12);
        NEAR(c.B + c.b1 + c.b2 + c.b3, 1.0, 1e-14);  // unit DC gain
    }
    // FIR radius floor(3 sigma + 1/2) and taps summing to 1 (reference src/fx/gaussian.cpp:24)
    CHECK(fir_radius(2.0) == 6 && fir_radius(8.0) == 24 && fir_radius(0.65) == 2);
    double sum = 0.0;
    for (float t : fir_taps(8.0)) sum += t;
    NEAR(sum, 1.0, 1e-6);
}

static void geometry_kats() {
    const GeometryMap id = geometry_map(0.0, 0.0, 1.0, 1.0, 0.0, 0, false, false, 600, 400);
    CHECK(id.identity && id.dst_w == 600 && id.dst_h == 400);
    const GeometryMap q1 = geometry_map(0.0, 0.0, 1.0, 1.0, 0.0, 1, false, false, 600, 400);
    CHECK(!q1.identity && q1.dst_w == 400 && q1.dst_h == 600);
    const GeometryMap crop = geometry_map(0.1, 0.2, 0.5, 0.25, 0.0, 0, false, false, 600, 400);
    CHECK(crop.dst_w == 300 && crop.dst_h == 100);
    // unrotated crop: output (0, 0) samples the crop's first pixel centre
    NEAR(crop.cx + crop.col_x[0] + crop.row_x[0], 0.1 * 600.0, 1e-12);
    NEAR(crop.cy + crop.col_y[0] + crop.row_y[0], 0.2 * 400.0, 1e-12);
}

static void stock_kats(const std::string& dir, const std::string& paper_id) {
    std::shared_ptr<res::Folder> f = res::Folder::open(dir);
    std::shared_ptr<const res::Stock> paper = f->stock(paper_id);
    const std::vector<double> y = paper_curves(*paper);
    const size_t K = paper->points();
    CHECK(y.size() == 3 * K);
    for (size_t c = 0; c < 3; ++c) {
        bool mono = true;
        for (size_t i = 1; i < K; ++i) mono = mono && y[3 * i + c] >= y[3 * (i - 1) + c];
        CHECK(mono);
        double amp = 0.0;
        for (size_t l = 0; l < paper->curve_model->amplitudes.cols; ++l) amp += paper->curve_model->amplitudes.at(c, l);
        CHECK(y[3 * (K - 1) + c] <= amp + 1e-12 && y[3 * (K - 1) + c] > 0.9 * amp);
        CHECK(y[c] < 0.01 * amp);
    }
    // a zero-density, base-free medium scans to Y = 1 under its viewing light (a little below
    // after the unmeasured bands drop out)
    const std::vector<double> cmf = f->constants->numbers("colour/cmfs_1931_2deg");
    const std::optional<Spectrum> view = res::light(*f->constants, paper->info.viewing_illuminant);
    CHECK(view.has_value());
    if (!view) return;
    const ScanSpectra s = scan_spectra(*paper, *view, cmf);
    double ysum = 0.0;
    for (size_t l = 0; l < 81; ++l) ysum += s.wt[3 * l + 1];
    CHECK(ysum <= 1.0 + 1e-14 && ysum > 0.99);
    NEAR(s.white[1], 1.0, 1e-14);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: setup_kats <resources folder> <paper stock id>\n");
        return 2;
    }
    try {
        mitchell_kats();
        exposure_kats();
        filter_kats();
        yvv_kats();
        geometry_kats();
        stock_kats(argv[1], argv[2]);
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf("%d failed\n", g_failed);
    return g_failed ? 1 : 0;
}
