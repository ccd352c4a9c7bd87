// mr_capi.cpp -- the C ABI of libmantaray_hip.so (include/mantaray_capi.h):
// field handles, their per-device tables, and the integration entry points.
//
// Citations are to files of mines-oceanography/mantaray 0.1.0.
#include "mantaray_capi.h"
#include "mr_guard.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mr_kernels.h"
#include "mr_netcdf3.h"

using namespace mr;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define MR_HIP_TRY(expr)                                                             \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(e_ == hipErrorOutOfMemory ? MR_ERR_OUT_OF_MEMORY : MR_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

bool is_bathy(int32_t k) { return k >= MR_BATHY_CONSTANT && k <= MR_BATHY_GRID; }
bool is_current(int32_t k) { return k == MR_CURRENT_CONSTANT || k == MR_CURRENT_GRID; }

// device allocations of one field on one device
struct DevTables {
    std::vector<void *> allocs;
    void *view = nullptr;          // BathyView or CurrentView in device memory
    BathyPtrs bp{};                // device pointers of a bathymetry
    CurrentPtrs cp{};              // device pointers of a current
    // fp64-interpolation mode, built on first use
    void *view64 = nullptr;        // Grid64View in device memory
    Grid64Ptrs p64{};
    const double *va = nullptr;    // the field's own values: depth, or u
    const double *vb = nullptr;    // v
    const SCell *scells = nullptr; // combined records with a current on shared cells
    ~DevTables()
    {
        for (void *p : allocs) (void)hipFree(p);
    }
};

// restores the caller's current device on scope exit
struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) return;
        ok = dev < 0 || hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <class T>
int dev_upload(DevTables &t, const T *host, size_t n, T **out)
{
    *out = nullptr;
    if (n == 0) return MR_OK;
    void *p = nullptr;
    MR_HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    t.allocs.push_back(p);
    MR_HIP_TRY(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<T *>(p);
    return MR_OK;
}

template <class T>
int dev_alloc(DevTables &t, size_t n, T **out)
{
    *out = nullptr;
    if (n == 0) return MR_OK;
    void *p = nullptr;
    MR_HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    t.allocs.push_back(p);
    *out = static_cast<T *>(p);
    return MR_OK;
}

// first cell (row-major, stride nx-1) whose largest corner index
// nx*(j+1) + i + 1 is past an array of len values
int64_t first_oob_cell(int64_t nx, int64_t ny, int64_t len)
{
    const int64_t ncx = nx - 1;
    for (int64_t j = 0; j < ny - 1; j++) {
        const int64_t c0 = nx * (j + 1) + 1;          // corner C of cell (0, j)
        if (c0 >= len) return j * ncx;
        if (c0 + (ncx - 1) >= len) return j * ncx + (len - c0);
    }
    return ncx * (ny - 1);
}

uint64_t next_field_uid()
{
    static std::atomic<uint64_t> n{1};
    return n.fetch_add(1);
}

}  // namespace

namespace mr {
// the error reporting of this file, for the other host translation units
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace mr

struct mr_field {
    uint64_t uid = next_field_uid();   // distinct for every field the process creates
    int32_t kind = 0;
    float h = 0, h0 = 0, x0 = 0, y0 = 0, dhdx = 0, dhdy = 0;
    double u = 0, v = 0;
    int64_t an = 0;
    std::vector<float> arr;
    int64_t nx = 0, ny = 0;
    std::vector<float> xf, yf;      // bathymetry axes (f32)
    std::vector<double> xd, yd;     // current axes (f64)
    std::vector<double> a, b;       // depth, or u and v
    std::mutex mu;
    std::map<int, std::unique_ptr<DevTables>> dev;
    // a bathymetry's combined records with a current on shared cells, per
    // (device * 16 + record tile, current uid)
    std::map<std::pair<int, uint64_t>, std::unique_ptr<DevTables>> shared;
};

struct mr_batch {
    int device = 0;
    int bk = 0, ck = 0;
    bool shared_cells = true;           // every scene's depth and current share cells
    int64_t n_rays = 0, max_ray_steps = 0;
    std::unique_ptr<DevTables> mem;     // descriptors, ray -> scene map, queue head
    SceneDesc *scenes = nullptr;
    int32_t *ray_scene = nullptr;
    unsigned long long *next = nullptr;
};

namespace {

int validate_axes_f32(const std::vector<float> &x, const char *name)
{
    if (x.size() < 2)
        return fail(MR_ERR_UNSUPPORTED,
                    std::string("axis '") + name +
                        "' has fewer than 2 points (the reference panics on every query)");
    for (float v : x)
        if (!std::isfinite(v))
            return fail(MR_ERR_UNSUPPORTED, std::string("axis '") + name + "' is not finite");
    if (x[1] == x[0])
        return fail(MR_ERR_UNSUPPORTED, std::string("axis '") + name + "' has zero spacing");
    if (x.size() >= (1u << 24))
        return fail(MR_ERR_UNSUPPORTED, std::string("axis '") + name + "' is too long");
    return MR_OK;
}

int validate_axes_f64(const std::vector<double> &x, const char *name)
{
    if (x.size() < 2)
        return fail(MR_ERR_UNSUPPORTED,
                    std::string("axis '") + name +
                        "' has fewer than 2 points (the reference panics on every query)");
    for (double v : x)
        if (!std::isfinite(v))
            return fail(MR_ERR_UNSUPPORTED, std::string("axis '") + name + "' is not finite");
    if (x[1] == x[0])
        return fail(MR_ERR_UNSUPPORTED, std::string("axis '") + name + "' has zero spacing");
    if (x.size() >= (1u << 24))
        return fail(MR_ERR_UNSUPPORTED, std::string("axis '") + name + "' is too long");
    return MR_OK;
}

// grids the kernels index with 32-bit cell byte offsets
// record tiles along x and y for 2^tl-cell tiles
int64_t tiles_along(int64_t n, int tl) { return (n - 1 + (int64_t(1) << tl) - 1) >> tl; }

int validate_grid(int64_t nx, int64_t ny, int64_t rec_bytes, int tl)
{
    if (nx >= 2 && ny >= 2 &&
        (tiles_along(nx, tl) * tiles_along(ny, tl) << (2 * tl)) * rec_bytes >= (int64_t(1) << 32))
        return fail(MR_ERR_UNSUPPORTED, "grid too large for 32-bit cell offsets");
    return MR_OK;
}

// ---- cell thresholds -------------------------------------------------------
// The reference's cell along an axis is a monotone step function of the
// coordinate: the fractional index RN(RN(p - a0) / sp) is non-decreasing in
// p (each rounding is monotone), and the cell is its floor, clamped to n-2 on
// the last node.  T[i] = min{p : index(p) >= i} for i < n-1 and T[n-1] =
// min{p : index(p) > n-1} make cell i exactly [T[i], T[i+1]) and the valid
// range [T[0], T[n-1]).  Each T is found by bisection over the totally
// ordered floating-point values, evaluating the index with the reference's
// own operations (cartesian_netcdf3.rs:285, cartesian_current.rs:244).

uint32_t key32(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
float unkey32(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
uint64_t key64(double f)
{
    uint64_t b;
    std::memcpy(&b, &f, 8);
    return (b >> 63) ? ~b : (b | (uint64_t(1) << 63));
}
double unkey64(uint64_t k)
{
    uint64_t b = (k >> 63) ? (k & ~(uint64_t(1) << 63)) : ~k;
    double f;
    std::memcpy(&f, &b, 8);
    return f;
}

// f32 axis (bathymetry): thresholds T[0..n-1]
std::vector<float> thresholds32(const std::vector<float> &a)
{
    const size_t n = a.size();
    const float a0 = a[0], sp = std::fabs(a[1] - a[0]);
    auto index = [&](float p) { return (p - a0) / sp; };        // :285
    std::vector<float> t(n);
    for (size_t i = 0; i < n; i++) {
        const float lim = (float)i;
        const bool upper = i == n - 1;                         // index > n-1 errs
        auto pred = [&](float p) { return upper ? index(p) > lim : index(p) >= lim; };
        uint32_t lo = key32(-INFINITY), hi = key32(INFINITY);  // pred(lo) false, pred(hi) true
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (pred(unkey32(mid))) hi = mid; else lo = mid;
        }
        t[i] = unkey32(hi);
    }
    return t;
}

// f64 axis (current): thresholds T[0..n-1]
std::vector<double> thresholds64(const std::vector<double> &a)
{
    const size_t n = a.size();
    const double a0 = a[0], sp = std::fabs(a[1] - a[0]);
    auto index = [&](double p) { return (p - a0) / sp; };       // :244
    std::vector<double> t(n);
    for (size_t i = 0; i < n; i++) {
        const double lim = (double)i;
        const bool upper = i == n - 1;
        auto pred = [&](double p) { return upper ? index(p) > lim : index(p) >= lim; };
        uint64_t lo = key64(-INFINITY), hi = key64(INFINITY);
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (pred(unkey64(mid))) hi = mid; else lo = mid;
        }
        t[i] = unkey64(hi);
    }
    return t;
}

// per-cell axis data of the bathymetry: (a[i], a[i+1], T[i], T[i+1])
std::vector<float4> axis_cells32(const std::vector<float> &a)
{
    const std::vector<float> t = thresholds32(a);
    std::vector<float4> c(a.size() - 1);
    for (size_t i = 0; i + 1 < a.size(); i++) c[i] = make_float4(a[i], a[i + 1], t[i], t[i + 1]);
    return c;
}

// per-cell thresholds of the current: (T[i], T[i+1])
std::vector<double2> axis_cells64(const std::vector<double> &a)
{
    const std::vector<double> t = thresholds64(a);
    std::vector<double2> c(a.size() - 1);
    for (size_t i = 0; i + 1 < a.size(); i++) c[i] = make_double2(t[i], t[i + 1]);
    return c;
}

// per-cell fast thresholds of the fp64 mode: the f64 index's cell
// [T[i], T[i+1]) narrowed to the open node interval (a[i], a[i+1]), so that a
// target inside on both axes is in the cell and on none of its node lines
// (an interval that comes out empty holds no target: every target of that
// cell then takes the full query)
std::vector<double2> axis_fast64(const std::vector<double> &a)
{
    const std::vector<double> t = thresholds64(a);
    std::vector<double2> c(a.size() - 1);
    for (size_t i = 0; i + 1 < a.size(); i++)
        c[i] = make_double2(std::max(t[i], std::nextafter(a[i], INFINITY)),
                            std::min(t[i + 1], a[i + 1]));
    return c;
}

std::vector<float2> axis_pairs(const std::vector<float> &a)
{
    std::vector<float2> p(a.size() - 1);
    for (size_t i = 0; i + 1 < a.size(); i++) p[i] = make_float2(a[i], a[i + 1]);
    return p;
}

int build_bathy(const mr_field &f, DevTables &t, int tl)
{
    BathyView v;
    std::memset(&v, 0, sizeof v);
    v.kind = f.kind;
    if (f.kind == MR_BATHY_CONSTANT) {
        v.h = f.h;
    } else if (f.kind == MR_BATHY_SLOPE) {
        v.h0 = f.h0; v.x0 = f.x0; v.y0 = f.y0; v.dhdx = f.dhdx; v.dhdy = f.dhdy;
    } else if (f.kind == MR_BATHY_ARRAY) {
        if (f.an > 46340) return fail(MR_ERR_UNSUPPORTED, "ArrayDepth too large");
        float *d = nullptr;
        int rc = dev_upload(t, f.arr.data(), f.arr.size(), &d);
        if (rc) return rc;
        t.bp.arr = d;
        v.an = (int)f.an;
    } else {
        int rc = validate_grid(f.nx, f.ny, 32, tl);
        if (!rc) rc = validate_axes_f32(f.xf, "x");
        if (!rc) rc = validate_axes_f32(f.yf, "y");
        if (rc) return rc;
        v.nx = (int)f.nx;
        v.ny = (int)f.ny;
        v.ncx = (int)(f.nx - 1);
        v.ax0 = f.xf[0];
        v.ay0 = f.yf[0];
        v.nxm1 = (float)(f.nx - 1);                 // (array.len() - 1) as f32
        v.nym1 = (float)(f.ny - 1);
        v.inv_sx = 1.0 / (double)std::fabs(f.xf[1] - f.xf[0]);   // spacing in f32
        v.inv_sy = 1.0 / (double)std::fabs(f.yf[1] - f.yf[0]);
        v.pinv_sx = (float)v.inv_sx;
        v.pinv_sy = (float)v.inv_sy;
        v.pmax_x = (float)(f.nx - 2);
        v.pmax_y = (float)(f.ny - 2);
        std::vector<float2> xp = axis_pairs(f.xf), yp = axis_pairs(f.yf);
        std::vector<float4> xa = axis_cells32(f.xf), ya = axis_cells32(f.yf);
        float2 *dxp, *dyp;
        float4 *dxa, *dya;
        double *draw;
        rc = dev_upload(t, xp.data(), xp.size(), &dxp);
        if (!rc) rc = dev_upload(t, yp.data(), yp.size(), &dyp);
        if (!rc) rc = dev_upload(t, xa.data(), xa.size(), &dxa);
        if (!rc) rc = dev_upload(t, ya.data(), ya.size(), &dya);
        if (!rc) rc = dev_upload(t, f.a.data(), f.a.size(), &draw);
        if (rc) return rc;
        const int64_t ntx = tiles_along(f.nx, tl), nty = tiles_along(f.ny, tl);
        const int64_t ntile_cells = (ntx * nty) << (2 * tl);    // 2^tl x 2^tl-cell tiles
        const int ntx_tl = (int)(ntx << 4 | tl);
        BCell *cells;
        rc = dev_alloc(t, (size_t)ntile_cells, &cells);
        if (rc) return rc;
        BathyBuildArgs ba;
        ba.nx = f.nx; ba.ny = f.ny; ba.xp = dxp; ba.yp = dyp; ba.raw = draw;
        ba.oob_from = first_oob_cell(f.nx, f.ny, (int64_t)f.a.size());
        ba.xspace = (double)f.xf[1] - (double)f.xf[0];   // cartesian_netcdf3.rs:119-120
        ba.yspace = (double)f.yf[1] - (double)f.yf[0];
        ba.cells = cells;
        ba.ntx = ntx_tl;
        MR_HIP_TRY(launch_build_bathy(ba, 0));
        v.oob_from = (uint32_t)ba.oob_from;
        v.ntx = ntx_tl;
        t.bp.xa = dxa; t.bp.ya = dya; t.bp.cells = cells; t.bp.raw = draw;
    }
    BathyView *dv;
    int rc = dev_upload(t, &v, 1, &dv);
    if (rc) return rc;
    MR_HIP_TRY(hipDeviceSynchronize());
    t.view = dv;
    return MR_OK;
}

int build_current(const mr_field &f, DevTables &t, int tl)
{
    CurrentView v;
    std::memset(&v, 0, sizeof v);
    v.kind = f.kind == MR_CURRENT_CONSTANT ? CK_CONSTANT : CK_GRID;
    if (f.kind == MR_CURRENT_CONSTANT) {
        v.u = f.u; v.v = f.v;
    } else {
        int rc = validate_grid(f.nx, f.ny, 64, tl);
        if (!rc) rc = validate_axes_f64(f.xd, "x");
        if (!rc) rc = validate_axes_f64(f.yd, "y");
        if (rc) return rc;
        v.nx = (int)f.nx;
        v.ny = (int)f.ny;
        v.ncx = (int)(f.nx - 1);
        v.cx0 = f.xd[0];
        v.cy0 = f.yd[0];
        v.sx = std::fabs(f.xd[1] - f.xd[0]);      // cartesian_current.rs:242
        v.sy = std::fabs(f.yd[1] - f.yd[0]);
        v.inv_sx = 1.0 / v.sx;
        v.inv_sy = 1.0 / v.sy;
        v.nxm1 = (double)(f.nx - 1);
        v.nym1 = (double)(f.ny - 1);
        v.px0 = (float)v.cx0;
        v.py0 = (float)v.cy0;
        v.pinv_sx = (float)v.inv_sx;
        v.pinv_sy = (float)v.inv_sy;
        v.pmax_x = (float)(f.nx - 2);
        v.pmax_y = (float)(f.ny - 2);
        std::vector<float> x32(f.xd.size()), y32(f.yd.size());
        for (size_t i = 0; i < x32.size(); i++) x32[i] = (float)f.xd[i];   // :375-378
        for (size_t i = 0; i < y32.size(); i++) y32[i] = (float)f.yd[i];
        std::vector<float2> xp = axis_pairs(x32), yp = axis_pairs(y32);
        std::vector<double2> xt = axis_cells64(f.xd), yt = axis_cells64(f.yd);
        float2 *dxp, *dyp;
        double2 *dxt, *dyt;
        double *du, *dv;
        rc = dev_upload(t, xp.data(), xp.size(), &dxp);
        if (!rc) rc = dev_upload(t, yp.data(), yp.size(), &dyp);
        if (!rc) rc = dev_upload(t, xt.data(), xt.size(), &dxt);
        if (!rc) rc = dev_upload(t, yt.data(), yt.size(), &dyt);
        if (!rc) rc = dev_upload(t, f.a.data(), f.a.size(), &du);
        if (!rc) rc = dev_upload(t, f.b.data(), f.b.size(), &dv);
        if (rc) return rc;
        const int64_t ntx = tiles_along(f.nx, tl), nty = tiles_along(f.ny, tl);
        const int64_t ntile_cells = (ntx * nty) << (2 * tl);    // 2^tl x 2^tl-cell tiles
        const int ntx_tl = (int)(ntx << 4 | tl);
        CCell *cells;
        float2 *mtab;
        rc = dev_alloc(t, (size_t)ntile_cells, &cells);
        if (!rc) rc = dev_alloc(t, (size_t)ntile_cells, &mtab);
        if (rc) return rc;
        CurrentBuildArgs ca;
        ca.nx = f.nx; ca.ny = f.ny; ca.xp = dxp; ca.yp = dyp; ca.u = du; ca.v = dv;
        ca.oob_from = std::min(first_oob_cell(f.nx, f.ny, (int64_t)f.a.size()),
                               first_oob_cell(f.nx, f.ny, (int64_t)f.b.size()));
        ca.xspace = f.xd[1] - f.xd[0];               // cartesian_current.rs:512-513
        ca.yspace = f.yd[1] - f.yd[0];
        ca.cells = cells;
        ca.mtab = mtab;
        ca.ntx = ntx_tl;
        MR_HIP_TRY(launch_build_current(ca, 0));
        v.oob_from = (uint32_t)ca.oob_from;
        v.ntx = ntx_tl;
        t.cp.xp = dxp; t.cp.yp = dyp; t.cp.xt = dxt; t.cp.yt = dyt;
        t.cp.cells = cells; t.cp.mtab = mtab;
        t.cp.raw_u = du; t.cp.raw_v = dv;
    }
    CurrentView *dview;
    int rc = dev_upload(t, &v, 1, &dview);
    if (rc) return rc;
    MR_HIP_TRY(hipDeviceSynchronize());
    t.view = dview;
    return MR_OK;
}

// log2 of the edge of the default record tile (8 x 8 cells)
constexpr int kDefaultTile = 3;

// the device view of a field on the current device, built on first use, per
// record-tile shape for the gridded kinds (tl = log2 of the tile edge)
int field_tables(const mr_field *cf, int tl, const DevTables **out)
{
    mr_field *f = const_cast<mr_field *>(cf);
    int dev = 0;
    MR_HIP_TRY(hipGetDevice(&dev));
    const bool gridded = f->kind == MR_BATHY_GRID || f->kind == MR_CURRENT_GRID;
    const int key = dev * 16 + (gridded ? tl : 0);
    std::lock_guard<std::mutex> lk(f->mu);
    auto it = f->dev.find(key);
    if (it != f->dev.end()) {
        *out = it->second.get();
        return MR_OK;
    }
    std::unique_ptr<DevTables> t(new DevTables);
    int rc = is_bathy(f->kind) ? build_bathy(*f, *t, tl) : build_current(*f, *t, tl);
    if (rc) return rc;
    *out = t.get();
    f->dev[key] = std::move(t);
    return MR_OK;
}

// the fp64-mode view of a gridded or analytic field (mr_device64.h)
int build64(const mr_field &f, DevTables &t)
{
    Grid64View v;
    std::memset(&v, 0, sizeof v);
    const bool bathy = is_bathy(f.kind);
    v.kind = bathy ? f.kind : (f.kind == MR_CURRENT_CONSTANT ? CK_CONSTANT : CK_GRID);
    if (bathy) {
        v.h = (double)f.h;
        v.h0 = (double)f.h0; v.x0 = (double)f.x0; v.y0 = (double)f.y0;
        v.dhdx = (double)f.dhdx; v.dhdy = (double)f.dhdy;
        v.an = (int)f.an;
        t.p64.arr = t.bp.arr;
        t.va = t.bp.raw;
    } else {
        v.u = f.u; v.v = f.v;
        t.va = t.cp.raw_u;
        t.vb = t.cp.raw_v;
    }
    if (f.kind == MR_BATHY_GRID || f.kind == MR_CURRENT_GRID) {
        std::vector<double> ax(f.nx), ay(f.ny);
        for (int64_t i = 0; i < f.nx; i++)
            ax[(size_t)i] = bathy ? (double)f.xf[(size_t)i] : f.xd[(size_t)i];
        for (int64_t i = 0; i < f.ny; i++)
            ay[(size_t)i] = bathy ? (double)f.yf[(size_t)i] : f.yd[(size_t)i];
        v.nx = (int)f.nx; v.ny = (int)f.ny; v.ncx = (int)(f.nx - 1);
        v.a0x = ax[0]; v.a0y = ay[0];
        v.sx = std::fabs(ax[1] - ax[0]); v.sy = std::fabs(ay[1] - ay[0]);
        v.inv_sx = 1.0 / v.sx; v.inv_sy = 1.0 / v.sy;
        v.nxm1 = (double)(f.nx - 1); v.nym1 = (double)(f.ny - 1);
        v.rxs = 1.0 / (ax[1] - ax[0]); v.rys = 1.0 / (ay[1] - ay[0]);
        v.oob_from = (uint32_t)(bathy ? first_oob_cell(f.nx, f.ny, (int64_t)f.a.size())
                                      : std::min(first_oob_cell(f.nx, f.ny, (int64_t)f.a.size()),
                                                 first_oob_cell(f.nx, f.ny, (int64_t)f.b.size())));
        v.last_a = (uint32_t)std::max<int64_t>(0, (int64_t)f.a.size() - 1);
        v.last_b = (uint32_t)std::max<int64_t>(0, (int64_t)f.b.size() - 1);
        v.gx0 = (float)ax[0]; v.gy0 = (float)ay[0];
        v.ginv_x = (float)v.inv_sx; v.ginv_y = (float)v.inv_sy;
        v.gmax_x = (float)(f.nx - 2); v.gmax_y = (float)(f.ny - 2);
        // the fast thresholds of the f64 index (p - a0)/|a1 - a0| on these
        // axes, and the node pairs of each cell
        std::vector<double2> xt = axis_fast64(ax), yt = axis_fast64(ay);
        std::vector<double2> xp(ax.size() - 1), yp(ay.size() - 1);
        for (size_t i = 0; i + 1 < ax.size(); i++) xp[i] = make_double2(ax[i], ax[i + 1]);
        for (size_t i = 0; i + 1 < ay.size(); i++) yp[i] = make_double2(ay[i], ay[i + 1]);
        double *dax, *day;
        double2 *dm, *dxt, *dyt, *dxp, *dyp;
        int rc = dev_upload(t, ax.data(), ax.size(), &dax);
        if (!rc) rc = dev_upload(t, ay.data(), ay.size(), &day);
        if (!rc) rc = dev_upload(t, xt.data(), xt.size(), &dxt);
        if (!rc) rc = dev_upload(t, yt.data(), yt.size(), &dyt);
        if (!rc) rc = dev_upload(t, xp.data(), xp.size(), &dxp);
        if (!rc) rc = dev_upload(t, yp.data(), yp.size(), &dyp);
        if (!rc) rc = dev_alloc(t, (size_t)((f.nx - 1) * (f.ny - 1)), &dm);
        if (rc) return rc;
        M64BuildArgs ma;
        ma.nx = f.nx; ma.ny = f.ny; ma.ax = dax; ma.ay = day; ma.m = dm;
        MR_HIP_TRY(launch_build_m64(ma, 0));
        t.p64.xt = dxt; t.p64.yt = dyt; t.p64.xp = dxp; t.p64.yp = dyp; t.p64.m = dm;
    }
    Grid64View *dv;
    int rc = dev_upload(t, &v, 1, &dv);
    if (rc) return rc;
    MR_HIP_TRY(hipDeviceSynchronize());
    t.view64 = dv;
    return MR_OK;
}

// field_tables plus the fp64-mode view
int field_tables64(const mr_field *cf, const DevTables **out)
{
    const DevTables *t0;
    int rc = field_tables(cf, kDefaultTile, &t0);
    if (rc) return rc;
    mr_field *f = const_cast<mr_field *>(cf);
    std::lock_guard<std::mutex> lk(f->mu);
    DevTables *t = const_cast<DevTables *>(t0);
    if (!t->view64 && (rc = build64(*f, *t))) return rc;
    *out = t;
    return MR_OK;
}

// the fp64 tables of a field on the current device if they exist; builds
// nothing (for the entry point that must not allocate)
int existing_tables64(const mr_field *cf, const DevTables **out)
{
    mr_field *f = const_cast<mr_field *>(cf);
    int dev = 0;
    MR_HIP_TRY(hipGetDevice(&dev));
    const bool gridded = f->kind == MR_BATHY_GRID || f->kind == MR_CURRENT_GRID;
    const int key = dev * 16 + (gridded ? kDefaultTile : 0);
    std::lock_guard<std::mutex> lk(f->mu);
    auto it = f->dev.find(key);
    if (it == f->dev.end() || !it->second->view64)
        return fail(MR_ERR_INVALID_ARGUMENT,
                    "a field has no fp64 tables on this device: trace it with mr_trace_fp64 "
                    "on the device first (this entry point builds nothing)");
    *out = it->second.get();
    return MR_OK;
}

int check_pair(const mr_field *b, const mr_field *c)
{
    if (!b || !c) return fail(MR_ERR_INVALID_ARGUMENT, "null field handle");
    if (!is_bathy(b->kind)) return fail(MR_ERR_INVALID_ARGUMENT, "first field is not a bathymetry");
    if (!is_current(c->kind)) return fail(MR_ERR_INVALID_ARGUMENT, "second field is not a current");
    return MR_OK;
}

int resolve_options(const mr_options *opt, mr_options &o, int max_block = 256)
{
    mr_options_init(&o);
    if (opt) {
        // a caller built before record_tile existed passes the shorter size
        const int32_t old_size = (int32_t)offsetof(mr_options, record_tile);
        if (opt->struct_size < old_size)
            return fail(MR_ERR_INVALID_ARGUMENT, "mr_options.struct_size too small");
        std::memcpy(&o, opt, std::min<size_t>((size_t)opt->struct_size, sizeof(mr_options)));
        o.struct_size = (int32_t)sizeof(mr_options);
    }
    if (o.record_tile == -1) o.record_tile = kDefaultTile;
    if (o.record_tile < 0 || o.record_tile > 5)
        return fail(MR_ERR_INVALID_ARGUMENT, "record_tile must be -1 or 0..5");
    if (o.precision != MR_PREC_REFERENCE && o.precision != MR_PREC_FP64_INTERP)
        return fail(MR_ERR_INVALID_ARGUMENT, "unknown precision mode");
    if (o.math != MR_MATH_FAST && o.math != MR_MATH_LITERAL)
        return fail(MR_ERR_INVALID_ARGUMENT, "unknown math mode");
    if (o.solout != MR_SOLOUT_STATE && o.solout != MR_SOLOUT_NEXT_DERIV)
        return fail(MR_ERR_INVALID_ARGUMENT, "unknown solout variant");
    if (o.memory != MR_MEM_HOST && o.memory != MR_MEM_DEVICE)
        return fail(MR_ERR_INVALID_ARGUMENT, "unknown memory kind");
    if (o.block_size != 0 && (o.block_size < 64 || o.block_size > max_block || o.block_size % 64))
        return fail(MR_ERR_INVALID_ARGUMENT,
                    "block_size must be a multiple of 64 up to " + std::to_string(max_block));
    return MR_OK;
}

int bathy_kind_of(const mr_field *b) { return b->kind; }   // MR_BATHY_* == BK_*
int current_kind_of(const mr_field *c)
{
    return c->kind == MR_CURRENT_CONSTANT ? CK_CONSTANT : CK_GRID;
}

int tuned_block(int64_t n_rays)
{
    // spread the rays over every CU (one workgroup each while rays last),
    // whole waves per workgroup, at most one wave per SIMD per workgroup
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    const int64_t per_cu = (n_rays + cus - 1) / cus;
    const int64_t waves = std::min<int64_t>(4, std::max<int64_t>(1, (per_cu + 63) / 64));
    return (int)(64 * waves);
}

// mr_rhs_many in fp64 mode
int rhs64_many(const mr_field *bathy, const mr_field *current, int64_t n, const double *s,
               double *ds, const mr_options &o)
{
    const DevTables *bt, *ct;
    int rc;
    if ((rc = field_tables64(bathy, &bt)) || (rc = field_tables64(current, &ct))) return rc;
    Rhs64Args a;
    a.B = bt->view64;
    a.C = ct->view64;
    a.bp = bt->p64;
    a.cp = ct->p64;
    a.f.depth = bt->va;
    a.f.u = ct->va;
    a.f.v = ct->vb;
    a.n = n;
    if (o.memory == MR_MEM_DEVICE) {
        a.s = s; a.ds = ds;
        MR_HIP_TRY(launch_rhs64(a, bathy_kind_of(bathy), current_kind_of(current), o.math,
                                (hipStream_t)o.stream));
        return MR_OK;
    }
    DevTables tmp;
    double *d_s, *d_ds;
    if ((rc = dev_alloc(tmp, 4 * (size_t)n, &d_s)) || (rc = dev_alloc(tmp, 4 * (size_t)n, &d_ds)))
        return rc;
    if (n) MR_HIP_TRY(hipMemcpy(d_s, s, 4 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    a.s = d_s; a.ds = d_ds;
    MR_HIP_TRY(launch_rhs64(a, bathy_kind_of(bathy), current_kind_of(current), o.math, 0));
    MR_HIP_TRY(hipDeviceSynchronize());
    if (n) MR_HIP_TRY(hipMemcpy(ds, d_ds, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return MR_OK;
}

// the fp64-mode trace, all pointers device pointers, enqueued on stream;
// depth, u, v replace the fields' own values when not null; ckpt (with
// stride > 0) receives the checkpoints
int trace64_device(const mr_field *bathy, const mr_field *current, const double *depth,
                   const double *u, const double *v, int64_t n_rays, const double *x0,
                   const double *y0, const double *kx0, const double *ky0, double t0, int64_t n,
                   double dt, int32_t stride, const mr_options &o, hipStream_t stream,
                   double *ckpt, double *t_out, double *x_out, double *y_out, double *kx_out,
                   double *ky_out, int32_t *len_out, double *end_out)
{
    const DevTables *bt, *ct;
    int rc = field_tables64(bathy, &bt);
    if (!rc) rc = field_tables64(current, &ct);
    if (rc) return rc;
    if (t_out && n_rays == 0) MR_HIP_TRY(launch_fill_time(t0, dt, n, t_out, stream));
    Trace64Args a;
    std::memset(&a, 0, sizeof a);
    a.B = bt->view64;
    a.C = ct->view64;
    a.bp = bt->p64;
    a.cp = ct->p64;
    a.f.depth = depth ? depth : bt->va;
    a.f.u = u ? u : ct->va;
    a.f.v = v ? v : ct->vb;
    a.n_rays = n_rays;
    a.n_steps = n;
    a.t0 = t0;
    a.t_out = t_out;
    a.dt = dt;
    a.half = dt / 2.0;
    a.sixth = dt / 6.0;
    a.x0 = x0; a.y0 = y0; a.kx0 = kx0; a.ky0 = ky0;
    a.xo = x_out; a.yo = y_out; a.kxo = kx_out; a.kyo = ky_out;
    a.len_out = len_out;
    a.end_out = end_out;
    a.solout = o.solout;
    a.stride = stride;
    a.ckpt = ckpt;
    const int block = o.block_size ? o.block_size : tuned_block(n_rays);
    if ((n_rays + block - 1) / block >= (int64_t)INT32_MAX)
        return fail(MR_ERR_UNSUPPORTED, "too many rays for one launch");
    a.ray_blocks = (int)((n_rays + block - 1) / block);
    MR_HIP_TRY(launch_trace64(a, bathy_kind_of(bathy), current_kind_of(current), o.math, block,
                              stream));
    return MR_OK;
}

// true when a gridded depth and a gridded current have the same cells: the
// same node counts, and the current's axes rounded to f32 equal the depth's
// (the trace then predicts both records' cells with one test; only the
// prefetch uses this, never a result)
bool shared_cells(const mr_field *b, const mr_field *c)
{
    if (b->kind != MR_BATHY_GRID || c->kind != MR_CURRENT_GRID) return false;
    if (b->nx != c->nx || b->ny != c->ny || b->xf.size() != c->xd.size() ||
        b->yf.size() != c->yd.size())
        return false;
    for (size_t i = 0; i < b->xf.size(); i++)
        if ((float)c->xd[i] != b->xf[i]) return false;
    for (size_t i = 0; i < b->yf.size(); i++)
        if ((float)c->yd[i] != b->yf[i]) return false;
    return true;
}

// The fast region of each cell along one axis (SCell): the f32 targets p in
// [lo, hi] lie in cell i for the depth's f32 index rule (T32[i] <= p <
// T32[i+1]), for every f64 coordinate that rounds to p lie in cell i for the
// current's f64 index rule (T64[i] <= x < T64[i+1]), and are strictly
// between the cell's node coordinates a[i] and a[i+1] (no corner).  If
// RN32(x) >= lo > ceil32(T64[i]) then x > T64[i]; if RN32(x) <= hi <
// floor32(T64[i+1]) then x < T64[i+1] (rounding is monotone).  The three
// intervals need not meet (on axes whose spacing varies, the index rules,
// which use the first spacing, can map a cell's targets outside its nodes):
// an empty interval is stored as (-inf, -inf), the one form of an empty
// region that the device's region test (med3(p, lo, hi) - p) reads as
// empty; lo > hi would read as [hi, lo].
std::vector<float2> fast_bounds(const std::vector<float> &a32, const std::vector<double> &a64)
{
    const std::vector<float> t32 = thresholds32(a32);
    const std::vector<double> t64 = thresholds64(a64);
    auto up = [](float f) { return std::nextafter(f, INFINITY); };
    auto down = [](float f) { return std::nextafter(f, -INFINITY); };
    auto ceil32 = [&](double d) {
        float f = (float)d;
        return (double)f < d ? up(f) : f;
    };
    auto floor32 = [&](double d) {
        float f = (float)d;
        return (double)f > d ? down(f) : f;
    };
    std::vector<float2> b(a32.size() - 1);
    for (size_t i = 0; i + 1 < a32.size(); i++) {
        const float lo = std::max({t32[i], up(a32[i]), up(ceil32(t64[i]))});
        const float hi = std::min({down(t32[i + 1]), down(a32[i + 1]), down(floor32(t64[i + 1]))});
        b[i] = lo <= hi ? make_float2(lo, hi) : make_float2(-INFINITY, -INFINITY);
    }
    return b;
}

// the combined records of a depth and a current on shared cells (SCell) on
// the current device, built on first use and kept with the bathymetry
int shared_tables(const mr_field *cb, const mr_field *cc, int tl, const SCell **out)
{
    const DevTables *bt, *ct;
    int rc = field_tables(cb, tl, &bt);
    if (!rc) rc = field_tables(cc, tl, &ct);
    if (rc) return rc;
    mr_field *b = const_cast<mr_field *>(cb);
    int dev = 0;
    MR_HIP_TRY(hipGetDevice(&dev));
    const std::pair<int, uint64_t> key(dev * 16 + tl, cc->uid);
    std::lock_guard<std::mutex> lk(b->mu);
    auto it = b->shared.find(key);
    if (it != b->shared.end()) {
        *out = it->second->scells;
        return MR_OK;
    }
    std::unique_ptr<DevTables> t(new DevTables);
    std::vector<double> cx(cc->xd), cy(cc->yd);
    std::vector<float2> fx = fast_bounds(b->xf, cx), fy = fast_bounds(b->yf, cy);
    float2 *dfx, *dfy;
    if ((rc = dev_upload(*t, fx.data(), fx.size(), &dfx)) ||
        (rc = dev_upload(*t, fy.data(), fy.size(), &dfy)))
        return rc;
    const int64_t ntx = tiles_along(b->nx, tl), nty = tiles_along(b->ny, tl);
    SCell *cells;
    if ((rc = validate_grid(b->nx, b->ny, (int64_t)sizeof(SCell), tl)) ||
        (rc = dev_alloc(*t, (size_t)((ntx * nty) << (2 * tl)), &cells)))
        return rc;
    SharedBuildArgs a;
    a.nx = b->nx; a.ny = b->ny;
    // the depth's axis pairs and raw values, and the current's raw values,
    // are already on the device with the fields' own tables
    std::vector<float2> xp = axis_pairs(b->xf), yp = axis_pairs(b->yf);
    float2 *dxp, *dyp;
    if ((rc = dev_upload(*t, xp.data(), xp.size(), &dxp)) ||
        (rc = dev_upload(*t, yp.data(), yp.size(), &dyp)))
        return rc;
    a.xp = dxp; a.yp = dyp; a.fx = dfx; a.fy = dfy;
    a.depth = bt->bp.raw; a.u = ct->cp.raw_u; a.v = ct->cp.raw_v;
    a.oob_from = std::min({first_oob_cell(b->nx, b->ny, (int64_t)b->a.size()),
                           first_oob_cell(b->nx, b->ny, (int64_t)cc->a.size()),
                           first_oob_cell(b->nx, b->ny, (int64_t)cc->b.size())});
    a.bxspace = (double)b->xf[1] - (double)b->xf[0];      // cartesian_netcdf3.rs:119-120
    a.byspace = (double)b->yf[1] - (double)b->yf[0];
    a.cxspace = cc->xd[1] - cc->xd[0];                      // cartesian_current.rs:512-513
    a.cyspace = cc->yd[1] - cc->yd[0];
    a.ntx = (int)(ntx << 4 | tl);
    a.cells = cells;
    MR_HIP_TRY(launch_build_shared(a, 0));
    MR_HIP_TRY(hipDeviceSynchronize());
    t->scells = cells;
    *out = cells;
    b->shared[key] = std::move(t);
    return MR_OK;
}

// shared cells whose combined records the kernels can index with 32-bit
// byte offsets (grids up to about 5790 x 5790 cells)
bool use_shared(const mr_field *b, const mr_field *c, int tl)
{
    return shared_cells(b, c) &&
           (tiles_along(b->nx, tl) * tiles_along(b->ny, tl) << (2 * tl)) * (int64_t)sizeof(SCell) <
               (int64_t(1) << 32);
}

// the trace itself, all pointers device pointers, enqueued on stream
int trace_device(const mr_field *bathy, const mr_field *current, int64_t n_rays,
                 const double *x0, const double *y0, const double *kx0, const double *ky0,
                 double t0, int64_t n, double dt, const mr_options &o, hipStream_t stream,
                 double *t_out, double *x_out, double *y_out, double *kx_out, double *ky_out,
                 int32_t *len_out, double *end_out)
{
    if (o.precision == MR_PREC_FP64_INTERP)
        return trace64_device(bathy, current, nullptr, nullptr, nullptr, n_rays, x0, y0, kx0, ky0,
                              t0, n, dt, 0, o, stream, nullptr, t_out, x_out, y_out, kx_out,
                              ky_out, len_out, end_out);
    const DevTables *bt, *ct;
    int rc = field_tables(bathy, o.record_tile, &bt);
    if (!rc) rc = field_tables(current, o.record_tile, &ct);
    if (rc) return rc;
    if (t_out && n_rays == 0) MR_HIP_TRY(launch_fill_time(t0, dt, n, t_out, stream));
    TraceArgs a;
    a.B = bt->view;
    a.C = ct->view;
    a.bp = bt->bp;
    a.cp = ct->cp;
    a.n_rays = n_rays;
    a.n_steps = n;
    a.t0 = t0;
    a.t_out = t_out;
    a.dt = dt;
    a.half = dt / 2.0;                                 // Rk4::new half_step
    a.sixth = dt / 6.0;                                // Rk4::step (step_size / 6)
    a.x0 = x0; a.y0 = y0; a.kx0 = kx0; a.ky0 = ky0;
    a.xo = x_out; a.yo = y_out; a.kxo = kx_out; a.kyo = ky_out;
    a.len_out = len_out;
    a.end_out = end_out;
    a.solout = o.solout;
    a.shared_cells = use_shared(bathy, current, o.record_tile) ? 1 : 0;
    a.scells = nullptr;
    if (a.shared_cells && (rc = shared_tables(bathy, current, o.record_tile, &a.scells)))
        return rc;
    const int block = o.block_size ? o.block_size : tuned_block(n_rays);
    if ((n_rays + block - 1) / block >= (int64_t)INT32_MAX)
        return fail(MR_ERR_UNSUPPORTED, "too many rays for one launch");
    a.ray_blocks = (int)((n_rays + block - 1) / block);
    MR_HIP_TRY(launch_trace(a, bathy_kind_of(bathy), current_kind_of(current), o.math, block,
                            stream));
    return MR_OK;
}

// mr_adjoint_many and mr_adjoint_many_noalloc: with no_alloc the work buffer
// is this device's scratch array and no field table is built
int adjoint_run(const mr_field *bathy, const mr_field *current, const double *depth,
                const double *u, const double *v, int64_t n_rays, double t0, double t_end,
                double dt, int32_t stride, const mr_options *opt, const double *ckpt,
                const int32_t *len, const double *end, const double *g_end, const double *g_x,
                const double *g_y, const double *g_kx, const double *g_ky, void *work,
                int64_t work_bytes, bool no_alloc, double *g_x0, double *g_y0, double *g_kx0,
                double *g_ky0, double *g_depth, double *g_u, double *g_v)
{
    int rc = check_pair(bathy, current);
    if (rc) return rc;
    mr_options o;
    if ((rc = resolve_options(opt, o))) return rc;
    if (o.memory != MR_MEM_DEVICE)
        return fail(MR_ERR_INVALID_ARGUMENT, "mr_adjoint_many takes device memory (MR_MEM_DEVICE)");
    if (o.solout != MR_SOLOUT_STATE)
        return fail(MR_ERR_INVALID_ARGUMENT, "the adjoint supports MR_SOLOUT_STATE only");
    if (n_rays < 0 || stride <= 0) return fail(MR_ERR_INVALID_ARGUMENT, "invalid sizes");
    if (g_depth && bathy->kind != MR_BATHY_GRID)
        return fail(MR_ERR_INVALID_ARGUMENT,
                    "depth gradient requested of a non-gridded bathymetry");
    if ((g_u || g_v) && current->kind != MR_CURRENT_GRID)
        return fail(MR_ERR_INVALID_ARGUMENT, "current gradient requested of a non-gridded current");
    int64_t need = 0;
    if ((rc = mr_adjoint_workspace_size(n_rays, stride, &need))) return rc;
    if (no_alloc && need > MR_ADJOINT_SCRATCH_BYTES)
        return fail(MR_ERR_INVALID_ARGUMENT,
                    "the sweep needs " + std::to_string(need) + " bytes of work buffer (" +
                        std::to_string(n_rays) + " rays at stride " + std::to_string(stride) +
                        "); mr_adjoint_many_noalloc holds " +
                        std::to_string((long long)MR_ADJOINT_SCRATCH_BYTES));
    if (n_rays > 0 && (!ckpt || !len || !end || (!no_alloc && !work)))
        return fail(MR_ERR_INVALID_ARGUMENT, "null checkpoint, length, end or work buffer");
    if (!no_alloc && work_bytes < need)
        return fail(MR_ERR_INVALID_ARGUMENT, "work buffer too small");
    int64_t n;
    if ((rc = mr_num_steps(t0, t_end, dt, &n))) return rc;
    DeviceGuard g(o.device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");
    const DevTables *bt, *ct;
    if (no_alloc) {
        if ((rc = existing_tables64(bathy, &bt)) || (rc = existing_tables64(current, &ct)))
            return rc;
        MR_HIP_TRY(adjoint_scratch(&work));
    } else if ((rc = field_tables64(bathy, &bt)) || (rc = field_tables64(current, &ct))) {
        return rc;
    }
    // the sweep kernel by workgroups (64 rays each) per CU: four waves per
    // workgroup up to one per CU (latency-bound: recompute | two linearise |
    // chain), two up to two per CU (recompute | sweep), one beyond (each wave
    // recomputes and sweeps in turn, the least SIMD time once the launch fills
    // the GPU); block_size 64, 128 or 256 forces one, two or four
    int cus = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    const int64_t groups = (n_rays + 63) / 64;
    int waves = groups <= cus ? 4 : groups <= 2 * (int64_t)cus ? 2 : 1;
    if (o.block_size == 64 || o.block_size == 128 || o.block_size == 256)
        waves = o.block_size / 64;
    else if (o.block_size != 0)
        return fail(MR_ERR_INVALID_ARGUMENT,
                    "mr_adjoint_many: block_size must be 0, 64, 128 or 256");
    hipStream_t stream = (hipStream_t)o.stream;
    if (g_depth) MR_HIP_TRY(launch_fill(g_depth, (int64_t)bathy->a.size(), 0.0, stream));
    if (g_u) MR_HIP_TRY(launch_fill(g_u, (int64_t)current->a.size(), 0.0, stream));
    if (g_v) MR_HIP_TRY(launch_fill(g_v, (int64_t)current->b.size(), 0.0, stream));
    AdjointArgs a;
    std::memset(&a, 0, sizeof a);
    a.B = bt->view64;
    a.C = ct->view64;
    a.bp = bt->p64;
    a.cp = ct->p64;
    a.f.depth = depth ? depth : bt->va;
    a.f.u = u ? u : ct->va;
    a.f.v = v ? v : ct->vb;
    a.n_rays = n_rays;
    a.n_steps = n;
    a.dt = dt;
    a.half = dt / 2.0;
    a.sixth = dt / 6.0;
    a.stride = stride;
    a.ckpt = ckpt;
    a.len = len;
    a.end = end;
    a.g_end = g_end;
    a.g_x = g_x; a.g_y = g_y; a.g_kx = g_kx; a.g_ky = g_ky;
    a.work = static_cast<double *>(work);
    a.work_ld = n_rays;
    a.g_x0 = g_x0; a.g_y0 = g_y0; a.g_kx0 = g_kx0; a.g_ky0 = g_ky0;
    a.grads.depth = g_depth;
    a.grads.u = g_u;
    a.grads.v = g_v;
    MR_HIP_TRY(launch_adjoint(a, bathy_kind_of(bathy), current_kind_of(current), o.math, waves,
                              stream));
    return MR_OK;
}

}  // namespace

extern "C" {

void mr_options_init(mr_options *opt)
{
    if (!opt) return;
    std::memset(opt, 0, sizeof *opt);
    opt->struct_size = (int32_t)sizeof(mr_options);
    opt->precision = MR_PREC_REFERENCE;
    opt->math = MR_MATH_FAST;
    opt->solout = MR_SOLOUT_STATE;
    opt->memory = MR_MEM_HOST;
    opt->device = -1;
    opt->block_size = 0;
    opt->blocks_per_cu = 0;
    opt->stream = nullptr;
    opt->record_tile = -1;
}

const char *mr_last_error(void) { return g_err.c_str(); }

int mr_version(int32_t *major, int32_t *minor, int32_t *patch)
{
    MR_GUARD_BEGIN
    if (major) *major = MR_VERSION_MAJOR;
    if (minor) *minor = MR_VERSION_MINOR;
    if (patch) *patch = MR_VERSION_PATCH;
    return MR_OK;
    MR_GUARD_END
}

int mr_bathy_from_arrays(int64_t nx, int64_t ny, const float *x, const float *y,
                         int64_t depth_len, const double *depth, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (nx < 0 || ny < 0 || depth_len < 0 || (nx && !x) || (ny && !y) || (depth_len && !depth))
        return fail(MR_ERR_INVALID_ARGUMENT, "invalid array arguments");
    std::unique_ptr<mr_field> f(new mr_field);
    f->kind = MR_BATHY_GRID;
    f->nx = nx; f->ny = ny;
    f->xf.assign(x, x + nx);
    f->yf.assign(y, y + ny);
    f->a.assign(depth, depth + depth_len);
    *out = f.release();
    return MR_OK;
    MR_GUARD_END
}

int mr_current_from_arrays(int64_t nx, int64_t ny, const double *x, const double *y,
                           int64_t u_len, const double *u, int64_t v_len, const double *v,
                           mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (nx < 0 || ny < 0 || u_len < 0 || v_len < 0 || (nx && !x) || (ny && !y) ||
        (u_len && !u) || (v_len && !v))
        return fail(MR_ERR_INVALID_ARGUMENT, "invalid array arguments");
    std::unique_ptr<mr_field> f(new mr_field);
    f->kind = MR_CURRENT_GRID;
    f->nx = nx; f->ny = ny;
    f->xd.assign(x, x + nx);
    f->yd.assign(y, y + ny);
    f->a.assign(u, u + u_len);
    f->b.assign(v, v + v_len);
    *out = f.release();
    return MR_OK;
    MR_GUARD_END
}

int mr_bathy_open_nc3(const char *path, const char *xname, const char *yname,
                      const char *depth_name, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out || !path || !xname || !yname || !depth_name)
        return fail(MR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    NcFile nc;
    std::string msg;
    int rc = nc3_open(path, nc, msg);
    if (rc) return fail(rc == -2 ? MR_ERR_IO : MR_ERR_FORMAT, msg);
    std::vector<float> x, y;
    std::vector<double> d;
    // cartesian_netcdf3.rs:170-252: axes promoted to f32, depth to f64
    if (nc3_read_f32(nc, xname, x, msg) || nc3_read_f32(nc, yname, y, msg) ||
        nc3_read_f64(nc, depth_name, d, msg))
        return fail(MR_ERR_FORMAT, msg);
    return mr_bathy_from_arrays((int64_t)x.size(), (int64_t)y.size(), x.data(), y.data(),
                                (int64_t)d.size(), d.data(), out);
    MR_GUARD_END
}

int mr_current_open_nc3(const char *path, const char *xname, const char *yname,
                        const char *uname, const char *vname, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out || !path || !xname || !yname || !uname || !vname)
        return fail(MR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    NcFile nc;
    std::string msg;
    int rc = nc3_open(path, nc, msg);
    if (rc) return fail(rc == -2 ? MR_ERR_IO : MR_ERR_FORMAT, msg);
    std::vector<double> x, y, u, v;
    // cartesian_current.rs:67-205: everything promoted to f64
    if (nc3_read_f64(nc, xname, x, msg) || nc3_read_f64(nc, yname, y, msg) ||
        nc3_read_f64(nc, uname, u, msg) || nc3_read_f64(nc, vname, v, msg))
        return fail(MR_ERR_FORMAT, msg);
    return mr_current_from_arrays((int64_t)x.size(), (int64_t)y.size(), x.data(), y.data(),
                                  (int64_t)u.size(), u.data(), (int64_t)v.size(), v.data(), out);
    MR_GUARD_END
}

int mr_bathy_constant(float h, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    mr_field *f = new mr_field;
    f->kind = MR_BATHY_CONSTANT;
    f->h = h;
    *out = f;
    return MR_OK;
    MR_GUARD_END
}

int mr_bathy_slope(float h0, float x0, float y0, float dhdx, float dhdy, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    mr_field *f = new mr_field;
    f->kind = MR_BATHY_SLOPE;
    f->h0 = h0; f->x0 = x0; f->y0 = y0; f->dhdx = dhdx; f->dhdy = dhdy;
    *out = f;
    return MR_OK;
    MR_GUARD_END
}

int mr_bathy_array(int64_t n, const float *a, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (n < 0 || (n && !a)) return fail(MR_ERR_INVALID_ARGUMENT, "invalid array arguments");
    mr_field *f = new mr_field;
    f->kind = MR_BATHY_ARRAY;
    f->an = n;
    f->arr.assign(a, a + n * n);
    *out = f;
    return MR_OK;
    MR_GUARD_END
}

int mr_current_constant(double u, double v, mr_field **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    mr_field *f = new mr_field;
    f->kind = MR_CURRENT_CONSTANT;
    f->u = u; f->v = v;
    *out = f;
    return MR_OK;
    MR_GUARD_END
}

int mr_field_free(mr_field *f)
{
    MR_GUARD_BEGIN
    delete f;
    return MR_OK;
    MR_GUARD_END
}

int mr_field_kind(const mr_field *f, int32_t *kind)
{
    MR_GUARD_BEGIN
    if (!f || !kind) return fail(MR_ERR_INVALID_ARGUMENT, "null argument");
    *kind = f->kind;
    return MR_OK;
    MR_GUARD_END
}

int mr_field_grid_shape(const mr_field *f, int64_t *nx, int64_t *ny, int64_t *len_a,
                        int64_t *len_b)
{
    MR_GUARD_BEGIN
    if (!f) return fail(MR_ERR_INVALID_ARGUMENT, "null field");
    if (nx) *nx = f->nx;
    if (ny) *ny = f->ny;
    if (len_a) *len_a = (int64_t)f->a.size();
    if (len_b) *len_b = (int64_t)f->b.size();
    return MR_OK;
    MR_GUARD_END
}

int mr_field_copy_grid(const mr_field *f, double *x, double *y, double *a, double *b)
{
    MR_GUARD_BEGIN
    if (!f) return fail(MR_ERR_INVALID_ARGUMENT, "null field");
    if (f->kind != MR_BATHY_GRID && f->kind != MR_CURRENT_GRID)
        return fail(MR_ERR_INVALID_ARGUMENT, "not a gridded field");
    for (int64_t i = 0; i < f->nx; i++)
        if (x) x[i] = f->kind == MR_BATHY_GRID ? (double)f->xf[(size_t)i] : f->xd[(size_t)i];
    for (int64_t i = 0; i < f->ny; i++)
        if (y) y[i] = f->kind == MR_BATHY_GRID ? (double)f->yf[(size_t)i] : f->yd[(size_t)i];
    if (a) std::memcpy(a, f->a.data(), f->a.size() * sizeof(double));
    if (b) std::memcpy(b, f->b.data(), f->b.size() * sizeof(double));
    return MR_OK;
    MR_GUARD_END
}

int mr_field_upload(const mr_field *f, int32_t device)
{
    MR_GUARD_BEGIN
    if (!f) return fail(MR_ERR_INVALID_ARGUMENT, "null field");
    DeviceGuard g(device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");
    const DevTables *t;
    return field_tables(f, kDefaultTile, &t);
    MR_GUARD_END
}

int mr_num_steps(double t0, double t_end, double dt, int64_t *n)
{
    MR_GUARD_BEGIN
    if (!n) return fail(MR_ERR_INVALID_ARGUMENT, "null output");
    // Rk4::integrate: ((x_end - x) / step_size).ceil().to_usize().unwrap();
    // num-traits accepts -1 < q < 2^64 and truncates
    const double q = std::ceil((t_end - t0) / dt);
    if (!(q > -1.0 && q < 18446744073709551616.0))
        return fail(MR_ERR_STEP_COUNT, "step count ceil((t_end - t0)/dt) is not a valid usize");
    if (q >= 2147483647.0) return fail(MR_ERR_STEP_COUNT, "step count too large");
    *n = (int64_t)q;
    return MR_OK;
    MR_GUARD_END
}

int mr_trace_many(const mr_field *bathy, const mr_field *current, int64_t n_rays,
                  const double *x0, const double *y0, const double *kx0, const double *ky0,
                  double t0, double t_end, double dt, const mr_options *opt,
                  double *t_out, double *x_out, double *y_out, double *kx_out,
                  double *ky_out, int32_t *len_out, double *end_out)
{
    MR_GUARD_BEGIN
    int rc = check_pair(bathy, current);
    if (rc) return rc;
    mr_options o;
    if ((rc = resolve_options(opt, o))) return rc;
    if (n_rays < 0) return fail(MR_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n_rays > 0 && (!x0 || !y0 || !kx0 || !ky0))
        return fail(MR_ERR_INVALID_ARGUMENT, "null initial conditions");
    int64_t n;
    if ((rc = mr_num_steps(t0, t_end, dt, &n))) return rc;
    DeviceGuard g(o.device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");

    if (o.memory == MR_MEM_DEVICE) {
        return trace_device(bathy, current, n_rays, x0, y0, kx0, ky0, t0, n, dt, o,
                            (hipStream_t)o.stream, t_out, x_out, y_out, kx_out, ky_out,
                            len_out, end_out);
    }
    // host memory: stage through device buffers, synchronous
    DevTables tmp;
    const size_t nr = (size_t)n_rays, rows = (size_t)n + 1;
    double *d_ic = nullptr, *d_t = nullptr, *d_end = nullptr;
    double *d_out[4] = {nullptr, nullptr, nullptr, nullptr};
    int32_t *d_len = nullptr;
    double *h_out[4] = {x_out, y_out, kx_out, ky_out};
    if ((rc = dev_alloc(tmp, 4 * nr, &d_ic))) return rc;
    const double *h_ic[4] = {x0, y0, kx0, ky0};
    for (int q = 0; q < 4; q++)
        if (nr) MR_HIP_TRY(hipMemcpy(d_ic + q * nr, h_ic[q], nr * sizeof(double),
                                     hipMemcpyHostToDevice));
    if (t_out && (rc = dev_alloc(tmp, rows, &d_t))) return rc;
    for (int q = 0; q < 4; q++)
        if (h_out[q] && (rc = dev_alloc(tmp, rows * nr, &d_out[q]))) return rc;
    if (len_out && (rc = dev_alloc(tmp, nr, &d_len))) return rc;
    if (end_out && (rc = dev_alloc(tmp, 4 * nr, &d_end))) return rc;
    rc = trace_device(bathy, current, n_rays, d_ic, d_ic + nr, d_ic + 2 * nr, d_ic + 3 * nr,
                      t0, n, dt, o, 0, d_t, d_out[0], d_out[1], d_out[2], d_out[3], d_len,
                      d_end);
    if (rc) return rc;
    MR_HIP_TRY(hipDeviceSynchronize());
    if (t_out) MR_HIP_TRY(hipMemcpy(t_out, d_t, rows * sizeof(double), hipMemcpyDeviceToHost));
    for (int q = 0; q < 4; q++)
        if (h_out[q] && nr)
            MR_HIP_TRY(hipMemcpy(h_out[q], d_out[q], rows * nr * sizeof(double),
                                 hipMemcpyDeviceToHost));
    if (len_out && nr)
        MR_HIP_TRY(hipMemcpy(len_out, d_len, nr * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (end_out && nr)
        MR_HIP_TRY(hipMemcpy(end_out, d_end, 4 * nr * sizeof(double), hipMemcpyDeviceToHost));
    return MR_OK;
    MR_GUARD_END
}

int mr_single_ray(const mr_field *bathy, const mr_field *current, double x0, double y0,
                  double kx0, double ky0, double t0, double t_end, double dt,
                  const mr_options *opt, double *rows_out, int64_t *n_rows)
{
    MR_GUARD_BEGIN
    if (!rows_out || !n_rows) return fail(MR_ERR_INVALID_ARGUMENT, "null output");
    mr_options o;
    int rc = resolve_options(opt, o);
    if (rc) return rc;
    if (o.memory != MR_MEM_HOST)
        return fail(MR_ERR_INVALID_ARGUMENT, "mr_single_ray takes host memory");
    int64_t n;
    if ((rc = mr_num_steps(t0, t_end, dt, &n))) return rc;
    const size_t rows = (size_t)n + 1;
    std::vector<double> t(rows), x(rows), y(rows), kx(rows), ky(rows);
    int32_t len = 0;
    rc = mr_trace_many(bathy, current, 1, &x0, &y0, &kx0, &ky0, t0, t_end, dt, &o, t.data(),
                       x.data(), y.data(), kx.data(), ky.data(), &len, nullptr);
    if (rc) return rc;
    for (int32_t r = 0; r < len; r++) {
        rows_out[5 * r + 0] = t[(size_t)r];
        rows_out[5 * r + 1] = x[(size_t)r];
        rows_out[5 * r + 2] = y[(size_t)r];
        rows_out[5 * r + 3] = kx[(size_t)r];
        rows_out[5 * r + 4] = ky[(size_t)r];
    }
    *n_rows = len;
    return MR_OK;
    MR_GUARD_END
}

int mr_rhs_many(const mr_field *bathy, const mr_field *current, int64_t n, const double *s,
                double *ds, const mr_options *opt)
{
    MR_GUARD_BEGIN
    int rc = check_pair(bathy, current);
    if (rc) return rc;
    mr_options o;
    if ((rc = resolve_options(opt, o))) return rc;
    if (n < 0 || (n && (!s || !ds))) return fail(MR_ERR_INVALID_ARGUMENT, "invalid arrays");
    DeviceGuard g(o.device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");
    if (o.precision == MR_PREC_FP64_INTERP) return rhs64_many(bathy, current, n, s, ds, o);
    const DevTables *bt, *ct;
    if ((rc = field_tables(bathy, o.record_tile, &bt)) ||
        (rc = field_tables(current, o.record_tile, &ct)))
        return rc;
    RhsArgs a;
    a.B = bt->view;
    a.C = ct->view;
    a.bp = bt->bp;
    a.cp = ct->cp;
    a.n = n;
    if (o.memory == MR_MEM_DEVICE) {
        a.s = s; a.ds = ds;
        MR_HIP_TRY(launch_rhs(a, bathy_kind_of(bathy), current_kind_of(current), o.math,
                              (hipStream_t)o.stream));
        return MR_OK;
    }
    DevTables tmp;
    double *d_s, *d_ds;
    if ((rc = dev_alloc(tmp, 4 * (size_t)n, &d_s)) || (rc = dev_alloc(tmp, 4 * (size_t)n, &d_ds)))
        return rc;
    if (n) MR_HIP_TRY(hipMemcpy(d_s, s, 4 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    a.s = d_s; a.ds = d_ds;
    MR_HIP_TRY(launch_rhs(a, bathy_kind_of(bathy), current_kind_of(current), o.math, 0));
    MR_HIP_TRY(hipDeviceSynchronize());
    if (n) MR_HIP_TRY(hipMemcpy(ds, d_ds, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return MR_OK;
    MR_GUARD_END
}

int mr_batch_create(int32_t n_scenes, const mr_scene *scenes, const mr_options *opt,
                    mr_batch **out)
{
    MR_GUARD_BEGIN
    if (!out) return fail(MR_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (n_scenes < 0 || (n_scenes > 0 && !scenes))
        return fail(MR_ERR_INVALID_ARGUMENT, "invalid scene array");
    mr_options o;
    int rc = resolve_options(opt, o);
    if (rc) return rc;
    if (o.memory != MR_MEM_DEVICE)
        return fail(MR_ERR_INVALID_ARGUMENT, "batches take device memory (MR_MEM_DEVICE)");
    if (o.precision != MR_PREC_REFERENCE)
        return fail(MR_ERR_UNSUPPORTED, "batches run in MR_PREC_REFERENCE");
    DeviceGuard g(o.device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");
    std::unique_ptr<mr_batch> b(new mr_batch);
    MR_HIP_TRY(hipGetDevice(&b->device));
    b->mem.reset(new DevTables);
    std::vector<SceneDesc> desc((size_t)n_scenes);
    std::vector<int32_t> ray_scene;
    for (int32_t i = 0; i < n_scenes; i++) {
        const mr_scene &sc = scenes[i];
        if ((rc = check_pair(sc.bathy, sc.current))) return rc;
        const int bk = bathy_kind_of(sc.bathy), ck = current_kind_of(sc.current);
        if (i == 0) { b->bk = bk; b->ck = ck; }
        else if (bk != b->bk || ck != b->ck)
            return fail(MR_ERR_INVALID_ARGUMENT, "scenes of one batch must share field kinds");
        b->shared_cells = b->shared_cells && use_shared(sc.bathy, sc.current, o.record_tile);
        if (sc.n_rays < 0) return fail(MR_ERR_INVALID_ARGUMENT, "negative ray count");
        if (sc.n_rays > 0 && (!sc.x0 || !sc.y0 || !sc.kx0 || !sc.ky0 || !sc.len_out || !sc.end_out))
            return fail(MR_ERR_INVALID_ARGUMENT, "null scene arrays");
        int64_t n;
        if ((rc = mr_num_steps(sc.t0, sc.t_end, sc.dt, &n))) return rc;
        const DevTables *bt, *ct;
        if ((rc = field_tables(sc.bathy, o.record_tile, &bt)) ||
            (rc = field_tables(sc.current, o.record_tile, &ct)))
            return rc;
        SceneDesc &d = desc[(size_t)i];
        d.B = bt->view; d.C = ct->view; d.bp = bt->bp; d.cp = ct->cp;
        d.dt = sc.dt; d.half = sc.dt / 2.0; d.sixth = sc.dt / 6.0;   // Rk4::new, Rk4::step
        d.n_steps = n;
        d.ray0 = b->n_rays;
        d.x0 = sc.x0; d.y0 = sc.y0; d.kx0 = sc.kx0; d.ky0 = sc.ky0;
        d.len_out = sc.len_out; d.end_out = sc.end_out;
        d.scells = nullptr;
        b->n_rays += sc.n_rays;
        b->max_ray_steps += sc.n_rays * n;
        ray_scene.insert(ray_scene.end(), (size_t)sc.n_rays, i);
    }
    if (b->n_rays >= (int64_t(1) << 31)) return fail(MR_ERR_UNSUPPORTED, "too many rays");
    if (b->shared_cells)
        for (int32_t i = 0; i < n_scenes; i++)
            if ((rc = shared_tables(scenes[i].bathy, scenes[i].current, o.record_tile,
                                    &desc[(size_t)i].scells)))
                return rc;
    if ((rc = dev_upload(*b->mem, desc.data(), desc.size(), &b->scenes)) ||
        (rc = dev_upload(*b->mem, ray_scene.data(), ray_scene.size(), &b->ray_scene)) ||
        (rc = dev_alloc(*b->mem, 1, &b->next)))
        return rc;
    MR_HIP_TRY(hipDeviceSynchronize());
    *out = b.release();
    return MR_OK;
    MR_GUARD_END
}

int mr_batch_run(mr_batch *b, const mr_options *opt)
{
    MR_GUARD_BEGIN
    if (!b) return fail(MR_ERR_INVALID_ARGUMENT, "null batch");
    mr_options o;
    int rc = resolve_options(opt, o);
    if (rc) return rc;
    if (o.memory != MR_MEM_DEVICE)
        return fail(MR_ERR_INVALID_ARGUMENT, "batches take device memory (MR_MEM_DEVICE)");
    if (o.blocks_per_cu < 0) return fail(MR_ERR_INVALID_ARGUMENT, "negative blocks_per_cu");
    DeviceGuard g(b->device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");
    if (b->n_rays == 0) return MR_OK;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) !=
            hipSuccess || cus <= 0)
        cus = 256;
    const int block = o.block_size ? o.block_size : 256;
    const int per_cu = o.blocks_per_cu ? o.blocks_per_cu : 1;
    const int64_t needed = (b->n_rays + block - 1) / block;
    const int blocks = (int)std::min<int64_t>(needed, (int64_t)cus * per_cu);
    hipStream_t stream = (hipStream_t)o.stream;
    BatchArgs a;
    a.scenes = b->scenes;
    a.ray_scene = b->ray_scene;
    a.n_rays = b->n_rays;
    a.next = b->next;
    a.solout = o.solout;
    a.shared_cells = b->shared_cells ? 1 : 0;
    MR_HIP_TRY(launch_clear_counter(b->next, stream));
    MR_HIP_TRY(launch_batch(a, b->bk, b->ck, o.math, block, blocks, stream));
    return MR_OK;
    MR_GUARD_END
}

int mr_batch_info(const mr_batch *b, int64_t *n_rays, int64_t *max_ray_steps)
{
    MR_GUARD_BEGIN
    if (!b) return fail(MR_ERR_INVALID_ARGUMENT, "null batch");
    if (n_rays) *n_rays = b->n_rays;
    if (max_ray_steps) *max_ray_steps = b->max_ray_steps;
    return MR_OK;
    MR_GUARD_END
}

int mr_batch_free(mr_batch *b)
{
    MR_GUARD_BEGIN
    if (b) {
        DeviceGuard g(b->device);
        delete b;
    }
    return MR_OK;
    MR_GUARD_END
}

int mr_checkpoint_count(double t0, double t_end, double dt, int32_t stride, int64_t *count)
{
    MR_GUARD_BEGIN
    if (!count) return fail(MR_ERR_INVALID_ARGUMENT, "null output");
    if (stride <= 0) return fail(MR_ERR_INVALID_ARGUMENT, "checkpoint stride must be positive");
    int64_t n;
    int rc = mr_num_steps(t0, t_end, dt, &n);
    if (rc) return rc;
    *count = std::max<int64_t>(1, (n + stride - 1) / stride);
    return MR_OK;
    MR_GUARD_END
}

int mr_adjoint_workspace_size(int64_t n_rays, int32_t stride, int64_t *bytes)
{
    MR_GUARD_BEGIN
    if (!bytes) return fail(MR_ERR_INVALID_ARGUMENT, "null output");
    if (n_rays < 0 || stride <= 0) return fail(MR_ERR_INVALID_ARGUMENT, "invalid sizes");
    // two halves (the recompute and the reverse sweep alternate between
    // them), each stride steps x 16 doubles x n_rays
    const int64_t per_ray = 2 * (int64_t)stride * 16 * (int64_t)sizeof(double);
    if (n_rays > INT64_MAX / per_ray)
        return fail(MR_ERR_INVALID_ARGUMENT, "work buffer size does not fit in 64 bits");
    *bytes = per_ray * n_rays;
    return MR_OK;
    MR_GUARD_END
}

int mr_trace_fp64(const mr_field *bathy, const mr_field *current, const double *depth,
                  const double *u, const double *v, int64_t n_rays, const double *x0,
                  const double *y0, const double *kx0, const double *ky0, double t0,
                  double t_end, double dt, int32_t stride, const mr_options *opt, double *ckpt,
                  double *t_out, double *x_out, double *y_out, double *kx_out, double *ky_out,
                  int32_t *len_out, double *end_out)
{
    MR_GUARD_BEGIN
    int rc = check_pair(bathy, current);
    if (rc) return rc;
    mr_options o;
    if ((rc = resolve_options(opt, o))) return rc;
    if (o.memory != MR_MEM_DEVICE)
        return fail(MR_ERR_INVALID_ARGUMENT, "mr_trace_fp64 takes device memory (MR_MEM_DEVICE)");
    if (n_rays < 0) return fail(MR_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n_rays > 0 && (!x0 || !y0 || !kx0 || !ky0))
        return fail(MR_ERR_INVALID_ARGUMENT, "null initial conditions");
    if (ckpt && stride <= 0)
        return fail(MR_ERR_INVALID_ARGUMENT, "checkpoint stride must be positive");
    int64_t n;
    if ((rc = mr_num_steps(t0, t_end, dt, &n))) return rc;
    DeviceGuard g(o.device);
    if (!g.ok) return fail(MR_ERR_DEVICE, "cannot select device");
    return trace64_device(bathy, current, depth, u, v, n_rays, x0, y0, kx0, ky0, t0, n, dt,
                          ckpt ? stride : 0, o, (hipStream_t)o.stream, ckpt, t_out, x_out, y_out,
                          kx_out, ky_out, len_out, end_out);
    MR_GUARD_END
}

int mr_adjoint_many(const mr_field *bathy, const mr_field *current, const double *depth,
                    const double *u, const double *v, int64_t n_rays, double t0, double t_end,
                    double dt, int32_t stride, const mr_options *opt, const double *ckpt,
                    const int32_t *len, const double *end, const double *g_end,
                    const double *g_x, const double *g_y, const double *g_kx, const double *g_ky,
                    void *work, int64_t work_bytes, double *g_x0, double *g_y0, double *g_kx0,
                    double *g_ky0, double *g_depth, double *g_u, double *g_v)
{
    MR_GUARD_BEGIN
    return adjoint_run(bathy, current, depth, u, v, n_rays, t0, t_end, dt, stride, opt, ckpt, len,
                       end, g_end, g_x, g_y, g_kx, g_ky, work, work_bytes, false, g_x0, g_y0,
                       g_kx0, g_ky0, g_depth, g_u, g_v);
    MR_GUARD_END
}

int mr_adjoint_many_noalloc(const mr_field *bathy, const mr_field *current, const double *depth,
                            const double *u, const double *v, int64_t n_rays, double t0,
                            double t_end, double dt, int32_t stride, const mr_options *opt,
                            const double *ckpt, const int32_t *len, const double *end,
                            const double *g_end, const double *g_x, const double *g_y,
                            const double *g_kx, const double *g_ky, double *g_x0, double *g_y0,
                            double *g_kx0, double *g_ky0, double *g_depth, double *g_u,
                            double *g_v)
{
    MR_GUARD_BEGIN
    return adjoint_run(bathy, current, depth, u, v, n_rays, t0, t_end, dt, stride, opt, ckpt, len,
                       end, g_end, g_x, g_y, g_kx, g_ky, nullptr, 0, true, g_x0, g_y0, g_kx0,
                       g_ky0, g_depth, g_u, g_v);
    MR_GUARD_END
}

}  // extern "C"
