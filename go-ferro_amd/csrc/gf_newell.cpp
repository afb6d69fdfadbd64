// gf_newell.cpp — host construction of the real-space demagnetisation
// tensor kernels (Newell, Williams & Dunlop 1993) exactly as the reference
// evaluates them (magnetic_initialize.cpp:1176-1557).
//
// Why on the host, in this exact form: the reference evaluates Newell's
// closed forms with log(r - z)-style terms whose second differences cancel
// catastrophically at large offsets, so its kernel carries roundoff noise of
// ~1e-10 of the peak (and the spectra are not exactly real).  Matching the
// reference to 1e-12 therefore needs the same values, bit for bit: the same
// branch structure, operation order and glibc libm, no FMA contraction
// (this file is compiled with -ffp-contract=off).  The work is split over
// host threads by x offset; each offset is evaluated independently, so the
// thread count does not change any value.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace gf {

namespace {

const double kPiN = 3.14159265358979323846;

// Newell f with the reference's special cases for vanishing coordinates
// (magnetic_initialize.cpp:1427-1476).
double nf(double x, double y, double z) {
    const double tiny = 1.0e-33;
    const double u = x * x, v = y * y, w = z * z;
    const double r = std::sqrt(u + v + w);
    const bool bx = std::fabs(x) > tiny, by = std::fabs(y) > tiny, bz = std::fabs(z) > tiny;
    const bool sx = std::fabs(x) < tiny, sy = std::fabs(y) < tiny, sz = std::fabs(z) < tiny;
    if (bx && by && bz)
        return z * (u - v) * std::log(r - z) / 2. + y * (u - w) * std::log(r - y) / 2. -
               x * y * z * std::atan(y * z / (x * r)) + (2. * u - v - w) * r / 6.;
    if (sx && by && bz) return -(z * v / 2.) * std::log(r - z) - (y * w / 2.) * std::log(r - y) - r * r * r / 6.;
    if (bx && sy && bz) return (z * u / 2.) * std::log(r - z) + (2. * u - w) * r / 6.;
    if (bx && by && sz) return (y * u / 2.) * std::log(r - y) + (2. * u - v) * r / 6.;
    if (sx && sy && bz) return -std::fabs(z) * std::fabs(z) * std::fabs(z) / 6.;
    if (sx && by && sz) return -std::fabs(y) * std::fabs(y) * std::fabs(y) / 6.;
    if (bx && sy && sz) return std::fabs(x) * std::fabs(x) * std::fabs(x) / 3.;
    return 0.;
}

// Newell g (magnetic_initialize.cpp:1507-1557).
double ng(double x, double y, double z) {
    const double tiny = 1.0e-33;
    const double u = x * x, v = y * y, w = z * z;
    const double r = std::sqrt(u + v + w);
    if (std::fabs(x * y * z) > tiny)
        return (y * y * y / 6. - y * w / 2.) * std::log(r + x) + (x / 6.) * (3. * w - u) * std::log(r - y) +
               x * y * r / 3. + x * y * z * std::log(r - z) + (v * z / 2.) * std::atan(x * z / (r * y)) +
               (z * z * z / 6.) * std::atan(x * y / (r * z)) + (u * z / 2.) * std::atan(y * z / (r * x));
    if (std::fabs(x) < tiny && std::fabs(y * z) > tiny) return (y * y * y / 6. - y * w / 2.) * std::log(r);
    if (std::fabs(x * z) > tiny && std::fabs(y) < tiny) return (x * w / 2. - x * x * x / 6.) * std::log(r);
    if (std::fabs(x * y) > tiny && std::fabs(z) < tiny)
        return ((y * y * y) / 6.) * std::log(r + x) - (x * x * x / 6.) * std::log(r - y) + x * y * r / 3.;
    if ((std::fabs(x) + std::fabs(y)) < tiny && std::fabs(z) > tiny) return 0.;
    if ((std::fabs(x) + std::fabs(z)) < tiny && std::fabs(y) > tiny) return (y * y * y) / 12. * std::log(v);
    if (std::fabs(x) > tiny && (std::fabs(y) + std::fabs(z)) < tiny) return -((x * x * x) / 12.) * std::log(u);
    return 0.;
}

// second difference of f along y (magnetic_initialize.cpp:1419-1425)
double nf_y(double x, double y, double z, double dy) {
    const double ylo = y - dy, yhi = y + dy;
    return 2. * nf(x, y, z) - nf(x, ylo, z) - nf(x, yhi, z);
}

// diagonal element: second differences along x and z of nf_y
// (magnetic_initialize.cpp:1404-1417; evaluation order kept)
double diag(double x, double y, double z, double dx, double dy, double dz) {
    const double xhi = x + dx, xlo = x - dx, zhi = z + dz, zlo = z - dz;
    return 4. * nf_y(x, y, z, dy) - 2. * nf_y(x, y, zhi, dy) - 2. * nf_y(x, y, zlo, dy) - 2. * nf_y(xhi, y, z, dy) +
           nf_y(xhi, y, zhi, dy) + nf_y(xhi, y, zlo, dy) - 2. * nf_y(xlo, y, z, dy) + nf_y(xlo, y, zhi, dy) +
           nf_y(xlo, y, zlo, dy);
}

// g differenced in x and z around half-step y offsets
// (magnetic_initialize.cpp:1486-1505; evaluation order kept)
double ng_xz(double x, double y, double z, double dx, double dy, double dz) {
    const double xhi = x + dx, xlo = x - dx, yhi = y + dy / 2., ylo = y - dy / 2., zhi = z + dz, zlo = z - dz;
    return 4. * ng(x, yhi, z) - 4. * ng(x, ylo, z) - 2. * ng(x, yhi, zhi) + 2. * ng(x, ylo, zhi) -
           2. * ng(x, yhi, zlo) + 2. * ng(x, ylo, zlo) - 2. * ng(xlo, yhi, z) + 2. * ng(xlo, ylo, z) +
           ng(xlo, yhi, zhi) - ng(xlo, ylo, zhi) + ng(xlo, yhi, zlo) - ng(xlo, ylo, zlo) -
           2. * ng(xhi, yhi, z) + 2. * ng(xhi, ylo, z) + ng(xhi, yhi, zhi) - ng(xhi, ylo, zhi) +
           ng(xhi, yhi, zlo) - ng(xhi, ylo, zlo);
}

// off-diagonal element (magnetic_initialize.cpp:1478-1484)
double offdiag(double x, double y, double z, double dx, double dy, double dz) {
    const double yhi = y + dy / 2., ylo = y - dy / 2.;
    return ng_xz(x, yhi, z, dx, dy, dz) - ng_xz(x, ylo, z, dx, dy, dz);
}

}  // namespace

// out9: nine n-arrays in the order xx, yy, zz, xy, yz, zx, yx, zy, xz
// (the reference's Axyz, Ayzx, Azxy, Bxyz, Byzx, Bzxy, Byxz, Bzyx, Bxzy),
// offsets |d| <= n/2-1 placed at d mod n, divided by n.
void newell_real_kernels(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz, double* out9,
                         int threads) {
    const int64_t n = nx * ny * nz;
    std::memset(out9, 0, sizeof(double) * 9 * static_cast<size_t>(n));
    const double denom = 2. * (2. * kPiN);
    const double dn = static_cast<double>(n);
    auto work = [&](int64_t ilo, int64_t ihi) {
        for (int64_t i = ilo; i < ihi; i++)
            for (int64_t j = 1 - ny / 2; j < ny / 2; j++)
                for (int64_t k = 1 - nz / 2; k < nz / 2; k++) {
                    const double x = static_cast<double>(i) * dx;
                    const double y = static_cast<double>(j) * dy;
                    const double z = static_cast<double>(k) * dz;
                    const int64_t cell = (((i + nx) % nx) * ny + (j + ny) % ny) * nz + (k + nz) % nz;
                    out9[0 * n + cell] = diag(x, y, z, dx, dy, dz) / denom / dn;
                    out9[1 * n + cell] = diag(y, z, x, dy, dz, dx) / denom / dn;
                    out9[2 * n + cell] = diag(z, x, y, dz, dx, dy) / denom / dn;
                    out9[3 * n + cell] = offdiag(x, y, z, dx, dy, dz) / denom / dn;
                    out9[4 * n + cell] = offdiag(y, z, x, dy, dz, dx) / denom / dn;
                    out9[5 * n + cell] = offdiag(z, x, y, dz, dx, dy) / denom / dn;
                    out9[6 * n + cell] = offdiag(y, x, z, dy, dx, dz) / denom / dn;
                    out9[7 * n + cell] = offdiag(z, y, x, dz, dy, dx) / denom / dn;
                    out9[8 * n + cell] = offdiag(x, z, y, dx, dz, dy) / denom / dn;
                }
    };
    const int64_t lo = 1 - nx / 2, hi = nx / 2;
    const int64_t span = hi - lo;
    if (span <= 0) return;
    int t = threads < 1 ? 1 : threads;
    if (t > span) t = static_cast<int>(span);
    std::vector<std::thread> pool;
    for (int w = 0; w < t; w++) {
        const int64_t a = lo + span * w / t, b = lo + span * (w + 1) / t;
        pool.emplace_back(work, a, b);
    }
    for (auto& th : pool) th.join();
}

}  // namespace gf
