// CPU check of the brick-list builders' two-level test (insert.hpp:
// superbrick, member, member_hit, superbrick_hit).  For every brick grid the
// library configures (the eight shapes at several box sizes) and many plane
// normals -- random, axis-aligned, diagonal, and normals of planes that just
// touch a brick's expanded box -- it checks that
//   1. the 2x2x2 groups partition the brick grid (member() and the group's
//      member list agree, every brick in exactly one group);
//   2. member_hit plus the group's disk filter is the per-brick test:
//      the plane through the origin with normal n passes within the
//      brick's expanded box (half extents (B - 1) / 2 + margin, margin 0.52
//      for owner-computes cores, 2.52 at kx = 0, 1.86 otherwise), and that
//      box reaches into the sample disk of radius NX -- restated here in
//      double precision from the geometry, with a tolerance band for the
//      kernels' fp32 rounding;
//   3. a group whose box the plane misses holds no member the per-brick test
//      keeps (the group test never drops a pair), checked against the
//      kernels' own fp32 member test.
// Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "insert.hpp"

using namespace spbp;

static long fails = 0, checks = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        ++checks;                                          \
        if (!(cond)) {                                     \
            if (fails < 20) std::fprintf(stderr, __VA_ARGS__); \
            ++fails;                                       \
        }                                                  \
    } while (0)

// The library's brick shapes: LDS region edges and owner-computes flag.
static const int SHAPES[8][4] = {{16, 16, 8, 0}, {16, 8, 8, 0},   {8, 8, 8, 0},    {16, 12, 8, 0},
                                 {16, 16, 10, 0}, {16, 16, 12, 0}, {16, 16, 13, 0}, {16, 16, 13, 1}};

static BrickGrid grid(int nx, int shape) {
    const int own = SHAPES[shape][3], halo = own ? 4 : 0;
    BrickGrid g;
    g.bx = SHAPES[shape][0] - halo;
    g.by = SHAPES[shape][1] - halo;
    g.bz = SHAPES[shape][2] - halo;
    g.own = own;
    g.nbx = (nx + g.bx - 1) / g.bx;
    g.nby = (2 * nx - 1 + g.by - 1) / g.by;
    g.nbz = (2 * nx - 1 + g.bz - 1) / g.bz;
    g.nb = g.nbx * g.nby * g.nbz;
    return g;
}

// Signed distance margin of the per-brick test in double precision:
// (reach - |n.c|) for the plane and the disk slack r^2 - |q|^2.
struct Exact {
    double plane, disk;
};
static Exact exact_test(const BrickGrid& bg, int b, int nx, double nxv, double nyv, double nzv) {
    const int i = b % bg.nbx, j = (b / bg.nbx) % bg.nby, k = b / (bg.nbx * bg.nby);
    const double x0 = i * bg.bx, y0 = -nx + 1 + j * bg.by, z0 = -nx + 1 + k * bg.bz;
    const double m = bg.own ? (x0 == 0 ? 2.52 : 0.52) : 1.86;
    const double hx = 0.5 * (bg.bx - 1) + m, hy = 0.5 * (bg.by - 1) + m, hz = 0.5 * (bg.bz - 1) + m;
    const double cx = x0 + 0.5 * (bg.bx - 1), cy = y0 + 0.5 * (bg.by - 1), cz = z0 + 0.5 * (bg.bz - 1);
    const double reach = hx * std::fabs(nxv) + hy * std::fabs(nyv) + hz * std::fabs(nzv) + 0.02;
    const double qx = std::fmax(std::fabs(cx) - hx, 0.0), qy = std::fmax(std::fabs(cy) - hy, 0.0),
                 qz = std::fmax(std::fabs(cz) - hz, 0.0);
    const double r = nx + 0.02;
    return {reach - std::fabs(nxv * cx + nyv * cy + nzv * cz), r * r - (qx * qx + qy * qy + qz * qz)};
}

static void run(int nx, int shape, const std::vector<float4>& normals) {
    const BrickGrid bg = grid(nx, shape);
    const int nsx = (bg.nbx + 1) / 2, nsy = (bg.nby + 1) / 2, nsz = (bg.nbz + 1) / 2, ns = nsx * nsy * nsz;
    // 1. partition
    std::vector<int> seen(bg.nb, 0);
    std::vector<SuperBrick> groups(ns);
    for (int s = 0; s < ns; ++s) {
        groups[s] = superbrick(bg, s, nx);
        for (int d = 0; d < 8; ++d) {
            const int b = member(bg, s, d);
            if (b >= 0) {
                ++seen[b];
                // the group's list holds the member unless its box misses the disk
                const Exact e = exact_test(bg, b, nx, 1.0, 0.0, 0.0);
                CHECK(groups[s].b[d] == b || (groups[s].b[d] == -1 && e.disk < 1e-3),
                      "nx %d shape %d: group %d member %d is %d, brick %d\n", nx, shape, s, d, groups[s].b[d], b);
            } else {
                CHECK(groups[s].b[d] == -1, "nx %d shape %d: group %d member %d past the edge is %d\n", nx, shape, s,
                      d, groups[s].b[d]);
            }
        }
    }
    for (int b = 0; b < bg.nb; ++b) CHECK(seen[b] == 1, "nx %d shape %d: brick %d in %d groups\n", nx, shape, b, seen[b]);
    // 2. and 3.
    for (const float4& g : normals) {
        for (int s = 0; s < ns; ++s) {
            const SuperBrick& sb = groups[s];
            const bool gh = superbrick_hit(sb, nx, g);
            for (int d = 0; d < 8; ++d) {
                const int b = member(bg, s, d);
                if (b < 0) continue;
                const bool mh = sb.b[d] >= 0 && member_hit(sb, d, g);
                const Exact e = exact_test(bg, b, nx, g.x, g.y, g.z);
                const bool exact = e.plane >= 0.0 && e.disk >= 0.0;
                // fp32 rounding of |n.c| and of the reach: 1e-3 voxels is far above it
                if (std::fabs(e.plane) > 1e-3 && std::fabs(e.disk) > 1e-3)
                    CHECK(mh == exact, "nx %d shape %d: brick %d normal (%g %g %g): kernel %d, geometry %d\n", nx,
                          shape, b, g.x, g.y, g.z, (int)mh, (int)exact);
                CHECK(!mh || gh, "nx %d shape %d: group %d drops brick %d for normal (%g %g %g)\n", nx, shape, s, b,
                      g.x, g.y, g.z);
            }
        }
    }
}

int main() {
    std::mt19937_64 rng(2026);
    std::normal_distribution<double> nd;
    auto unit = [](double x, double y, double z) {
        const double r = std::sqrt(x * x + y * y + z * z);
        return make_float4((float)(x / r), (float)(y / r), (float)(z / r), 0.0f);
    };
    std::vector<float4> base;
    for (int a = 0; a < 3; ++a)
        for (int sgn : {-1, 1}) base.push_back(unit(a == 0 ? sgn : 0, a == 1 ? sgn : 0, a == 2 ? sgn : 0));
    base.push_back(unit(1, 1, 0));
    base.push_back(unit(1, -1, 0));
    base.push_back(unit(0, 1, 1));
    base.push_back(unit(1, 1, 1));
    base.push_back(unit(1, -1, 1));
    base.push_back(unit(1e-7, 1, 1e-7));
    for (int nx : {16, 24, 40, 64, 128}) {
        for (int shape = 0; shape < 8; ++shape) {
            std::vector<float4> normals = base;
            const int nrand = nx >= 128 ? 150 : 400;
            for (int k = 0; k < nrand; ++k) normals.push_back(unit(nd(rng), nd(rng), nd(rng)));
            // normals of planes through a corner of a brick's expanded box:
            // the tangent cases of the plane test
            const BrickGrid bg = grid(nx, shape);
            std::uniform_int_distribution<int> pick(0, bg.nb - 1);
            for (int k = 0; k < 100; ++k) {
                const int b = pick(rng);
                const int i = b % bg.nbx, j = (b / bg.nbx) % bg.nby, kk = b / (bg.nbx * bg.nby);
                const double m = bg.own ? (i == 0 ? 2.52 : 0.52) : 1.86;
                const double cx = i * bg.bx - m + (rng() & 1) * (bg.bx - 1 + 2 * m);
                const double cy = -nx + 1 + j * bg.by - m + (rng() & 1) * (bg.by - 1 + 2 * m);
                const double cz = -nx + 1 + kk * bg.bz - m + (rng() & 1) * (bg.bz - 1 + 2 * m);
                // a normal orthogonal to the corner: the plane passes through it
                double ux = nd(rng), uy = nd(rng), uz = nd(rng);
                const double cc = cx * cx + cy * cy + cz * cz;
                if (cc == 0.0) continue;
                const double t = (ux * cx + uy * cy + uz * cz) / cc;
                ux -= t * cx;
                uy -= t * cy;
                uz -= t * cz;
                if (ux * ux + uy * uy + uz * uz < 1e-12) continue;
                normals.push_back(unit(ux, uy, uz));
            }
            run(nx, shape, normals);
        }
    }
    std::printf("%ld checks, %ld failures\n", checks, fails);
    if (fails) return 1;
    std::printf("brick group checks passed\n");
    return 0;
}
