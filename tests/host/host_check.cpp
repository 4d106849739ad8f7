// Host-side checks of libspbp under AddressSanitizer and UndefinedBehavior-
// Sanitizer: every C-ABI entry point is called on the paths a host without a
// usable GPU reaches -- argument validation, NULL contexts and pointers,
// out-of-range sizes, the host-only table and rotation routines, context
// creation failing at the device, the fixed-size error buffers -- and the
// results are checked.  Exit status 0 when every check holds; any sanitizer
// report aborts the run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "spbp.h"

static int fails = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++fails;                                                             \
        }                                                                        \
    } while (0)

static bool msg_ok(const char* m) { return m && std::strlen(m) > 0 && std::strlen(m) < 4096; }

int main() {
    CHECK(msg_ok(spbp_version()));

    // KB tables (host arithmetic): TABI(0) = 1, FLTB = 4999/3, decreasing
    std::vector<float> tab(5000);
    float fltb = 0.f;
    for (int n2 : {8, 30, 128, 200, 512, 2048})
        for (int f : {0, SPBP_GEOM_F64}) {
            CHECK(spbp_kb_table(n2, f, tab.data(), &fltb) == SPBP_OK);
            CHECK(std::fabs(fltb - 4999.0f / 3.0f) < 1e-3f);
            CHECK(tab[0] > 0.999f && tab[0] < 1.001f && tab[4999] < tab[0]);
        }
    CHECK(spbp_kb_table(0, 0, tab.data(), &fltb) != SPBP_OK);
    CHECK(spbp_kb_table(128, 0, nullptr, &fltb) != SPBP_OK);

    // CANG rotations (host): orthonormal rows
    const int n = 257;
    std::vector<double> ang(3 * n);
    for (int i = 0; i < 3 * n; ++i) ang[i] = std::fmod(37.1 * i, 720.0) - 360.0;
    std::vector<float> dm(9 * n);
    for (int f : {0, SPBP_EXACT_PI, SPBP_GEOM_F64}) {
        CHECK(spbp_rotations(n, ang.data(), f, dm.data()) == SPBP_OK);
        for (int p = 0; p < n; ++p) {
            const float* r = &dm[9 * p];
            const double d01 = (double)r[0] * r[3] + (double)r[1] * r[4] + (double)r[2] * r[5];
            const double d00 = (double)r[0] * r[0] + (double)r[1] * r[1] + (double)r[2] * r[2];
            CHECK(std::fabs(d01) < 1e-5 && std::fabs(d00 - 1.0) < 1e-5);
        }
    }
    CHECK(spbp_rotations(-1, ang.data(), 0, dm.data()) != SPBP_OK);
    CHECK(spbp_rotations(3, nullptr, 0, dm.data()) != SPBP_OK);

    // context creation: bad options, then the device
    spbp_opts o;
    std::memset(&o, 0, sizeof o);
    spbp_ctx* c = nullptr;
    CHECK(spbp_create(nullptr, &c) == SPBP_ERR_ARG);
    o.nx = 2;
    CHECK(spbp_create(&o, &c) == SPBP_ERR_SIZE && c == nullptr);
    o.nx = 5000;
    CHECK(spbp_create(&o, &c) == SPBP_ERR_SIZE && c == nullptr);
    o.nx = 32;
    o.strategy = 9;
    CHECK(spbp_create(&o, &c) != SPBP_OK && c == nullptr);
    o.strategy = 0;
    o.device = 1 << 20;
    CHECK(spbp_create(&o, &c) == SPBP_ERR_DEVICE && c == nullptr);
    CHECK(msg_ok(spbp_last_error(nullptr)));
    o.device = 0;
    const int rc = spbp_create(&o, &c);
    CHECK(rc == SPBP_OK || rc == SPBP_ERR_DEVICE);
    if (rc == SPBP_OK) CHECK(spbp_destroy(c) == SPBP_OK);
    c = nullptr;

    // every context entry point with a NULL context fails cleanly
    float buf[64] = {0};
    double dbuf[64] = {0};
    long long ll[3];
    long l1, l2;
    int i1, i2, i3;
    CHECK(spbp_reset(nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_insert(nullptr, buf, SPBP_F32, 1, dbuf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_insert_dm(nullptr, buf, SPBP_F32, 1, buf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_insert_host(nullptr, buf, SPBP_F32, 1, dbuf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_insert_rb(nullptr, buf, SPBP_F32, 1, dbuf, buf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_insert_rb_host(nullptr, buf, SPBP_F32, 1, dbuf, buf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_finalize(nullptr, buf, nullptr, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_finalize_host(nullptr, buf, nullptr, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_export_partial(nullptr, -1, buf, buf, nullptr) != SPBP_OK);
    CHECK(spbp_import_partial(nullptr, -1, buf, buf, 1, nullptr) != SPBP_OK);
    CHECK(spbp_export_raw(nullptr, 0, buf, buf, nullptr) != SPBP_OK);
    CHECK(spbp_kernel_stats(nullptr, 0, dbuf, &l1, &l2) != SPBP_OK);
    CHECK(spbp_reset_stats(nullptr) != SPBP_OK);
    CHECK(spbp_threshold_stats(nullptr, &ll[0], &ll[1], &ll[2]) != SPBP_OK);
    CHECK(spbp_reference_weights(nullptr, -1, ll, 1, dbuf, nullptr) != SPBP_OK);
    CHECK(spbp_grid_elems(nullptr) == 0);
    CHECK(spbp_gather(nullptr, buf, 1, dbuf, nullptr, buf, nullptr) != SPBP_OK);
    CHECK(spbp_scatter(nullptr, buf, 1, dbuf, nullptr, buf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_gather_images(nullptr, buf, 1, dbuf, nullptr, buf, nullptr) != SPBP_OK);
    CHECK(spbp_scatter_images(nullptr, buf, 1, dbuf, nullptr, buf, nullptr) != SPBP_OK);
    CHECK(spbp_set_symmetries(nullptr, 1, dbuf) != SPBP_OK);
    CHECK(spbp_tune(nullptr, buf, SPBP_F32, 1, dbuf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_get_plan(nullptr, &i1, &i2, &i3, buf) != SPBP_OK);
    CHECK(spbp_set_plan(nullptr, SPBP_STRAT_BRICK, 0, 0) != SPBP_OK);
    CHECK(msg_ok(spbp_last_error(nullptr)));

    // batches: empty, oversized and NULL-context job lists
    spbp_job jobs[2];
    std::memset(jobs, 0, sizeof jobs);
    CHECK(spbp_insert_batch(nullptr, 1, nullptr) != SPBP_OK);
    CHECK(spbp_insert_batch(jobs, 0, nullptr) != SPBP_OK);
    CHECK(spbp_insert_batch(jobs, 1000, nullptr) != SPBP_OK);
    CHECK(spbp_insert_batch(jobs, 2, nullptr) != SPBP_OK);
    spbp_graph* g = nullptr;
    CHECK(spbp_graph_create(jobs, 2, 1, &g) != SPBP_OK && g == nullptr);
    CHECK(spbp_graph_create(jobs, 1, 1, nullptr) != SPBP_OK);
    CHECK(spbp_graph_launch(nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_graph_destroy(nullptr) == SPBP_OK);
    spbp_final_job fj[2];
    std::memset(fj, 0, sizeof fj);
    CHECK(spbp_finalize_batch(nullptr, 1, nullptr) != SPBP_OK);
    CHECK(spbp_finalize_batch(fj, 0, nullptr) != SPBP_OK);
    CHECK(spbp_finalize_batch(fj, 65, nullptr) != SPBP_OK);
    CHECK(spbp_finalize_batch(fj, 2, nullptr) != SPBP_OK);

    // context-free entry points: argument checks before any device work
    CHECK(spbp_rtsq(nullptr, SPBP_F32, 1, 16, buf, buf, nullptr) != SPBP_OK);
    CHECK(spbp_rtsq(buf, SPBP_F32, -1, 16, buf, buf, nullptr) != SPBP_OK);
    CHECK(spbp_rtsq(buf, 7, 1, 16, buf, buf, nullptr) != SPBP_OK);
    CHECK(spbp_rtsq(buf, SPBP_F32, 0, 16, buf, buf, nullptr) == SPBP_OK);
    CHECK(spbp_rtsq_host(nullptr, 1, 16, buf, buf) != SPBP_OK);
    CHECK(spbp_rtsq_host(buf, 0, 16, buf, buf) == SPBP_OK);
    CHECK(spbp_scatter_noalloc(800, 0, buf, 1, dbuf, nullptr, buf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_scatter_noalloc(2, 0, buf, 1, dbuf, nullptr, buf, nullptr, nullptr) != SPBP_OK);
    CHECK(spbp_scatter_noalloc(16, 0, nullptr, 1, dbuf, nullptr, buf, nullptr, nullptr) != SPBP_OK);
    double rate = 0.0;
    CHECK(spbp_calibrate(7, 0, &rate, nullptr) == SPBP_ERR_ARG);
    CHECK(spbp_calibrate(SPBP_CAL_LDS_U64, 0, nullptr, nullptr) == SPBP_ERR_ARG);
    CHECK(spbp_debug_fail_alloc(-1) == SPBP_OK);
    CHECK(msg_ok(spbp_last_error(nullptr)));

    if (fails) std::fprintf(stderr, "%d host checks failed\n", fails);
    else std::printf("host checks passed\n");
    return fails ? 1 : 0;
}
