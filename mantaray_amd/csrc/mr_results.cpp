// mr_results.cpp -- the reference's result writers, host side: RayResult's
// JSON (src/ray_result.rs:74-151, through serde_json) and the TSV writer
// output_or_append_to_tsv_file (src/ray.rs:218-243).
//
// Both print each f64 with the shortest decimal digits that read back to the
// same value; they differ in layout:
//   - serde_json writes f64 with the ryu crate's layout: decimal notation
//     with a ".0" for integral values while the decimal exponent kk (with
//     10^(kk-1) <= |v| < 10^kk) lies in (-5, 16], scientific "d.ddde-7"
//     otherwise, and null for NaN and infinities;
//   - Rust's Display for f64 (the TSV writer's "{}") never uses an exponent
//     and prints integral values without a fraction ("1", "-0", "NaN", "inf").
// The digits come from std::to_chars in shortest round-trip form.
#include "mantaray_capi.h"
#include "mr_guard.h"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

// shortest round-trip digits of |v| (finite, nonzero) and kk with
// 10^(kk-1) <= |v| < 10^kk
void shortest(double v, std::string &digits, int &kk)
{
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof buf, std::fabs(v), std::chars_format::scientific);
    *res.ptr = '\0';
    const char *e = std::strchr(buf, 'e');
    digits.clear();
    for (const char *p = buf; p < e; ++p)
        if (*p != '.') digits.push_back(*p);
    kk = std::atoi(e + 1) + 1;
}

// serde_json's f64 (the ryu crate's format64 layout)
void json_number(double v, std::string &out)
{
    if (!std::isfinite(v)) {
        out += "null";
        return;
    }
    if (std::signbit(v)) out.push_back('-');
    if (v == 0.0) {
        out += "0.0";
        return;
    }
    std::string d;
    int kk;
    shortest(v, d, kk);
    const int len = (int)d.size();
    const int k = kk - len;
    if (k >= 0 && kk <= 16) {                   // 1234e7 -> 12340000000.0
        out += d;
        out.append((size_t)k, '0');
        out += ".0";
    } else if (kk > 0 && kk <= 16) {            // 1234e-2 -> 12.34
        out.append(d, 0, (size_t)kk);
        out.push_back('.');
        out.append(d, (size_t)kk, std::string::npos);
    } else if (kk > -5 && kk <= 0) {            // 1234e-6 -> 0.001234
        out += "0.";
        out.append((size_t)-kk, '0');
        out += d;
    } else if (len == 1) {                      // 1e30
        out += d;
        out.push_back('e');
        out += std::to_string(kk - 1);
    } else {                                    // 1234e30 -> 1.234e33
        out.push_back(d[0]);
        out.push_back('.');
        out.append(d, 1, std::string::npos);
        out.push_back('e');
        out += std::to_string(kk - 1);
    }
}

// Rust's Display for f64: shortest digits, no exponent
void display_number(double v, std::string &out)
{
    if (std::isnan(v)) {
        out += "NaN";
        return;
    }
    if (std::signbit(v)) out.push_back('-');
    if (std::isinf(v)) {
        out += "inf";
        return;
    }
    if (v == 0.0) {
        out.push_back('0');
        return;
    }
    std::string d;
    int kk;
    shortest(v, d, kk);
    const int len = (int)d.size();
    if (kk <= 0) {
        out += "0.";
        out.append((size_t)-kk, '0');
        out += d;
    } else if (kk >= len) {
        out += d;
        out.append((size_t)(kk - len), '0');
    } else {
        out.append(d, 0, (size_t)kk);
        out.push_back('.');
        out.append(d, (size_t)kk, std::string::npos);
    }
}

}  // namespace

extern "C" {

int mr_ray_result_json(const double *t, const double *x, const double *y, const double *kx,
                       const double *ky, int64_t n_rows, int64_t stride, char *buf,
                       int64_t cap, int64_t *len)
{
    MR_GUARD_BEGIN
    if (!len || n_rows < 0 || stride < 1 || (n_rows > 0 && (!t || !x || !y || !kx || !ky)))
        return mr::set_error(MR_ERR_INVALID_ARGUMENT, "invalid arrays");
    // RayResult::from (ray_result.rs:124-151): rows up to the first with a
    // NaN in any component
    int64_t rows = 0;
    while (rows < n_rows) {
        const int64_t o = rows * stride;
        if (std::isnan(x[o]) || std::isnan(y[o]) || std::isnan(kx[o]) || std::isnan(ky[o])) break;
        ++rows;
    }
    std::string s = "{";
    const char *names[5] = {"t_vec", "x_vec", "y_vec", "kx_vec", "ky_vec"};
    const double *cols[5] = {t, x, y, kx, ky};
    for (int c = 0; c < 5; ++c) {
        if (c) s.push_back(',');
        s += '"';
        s += names[c];
        s += "\":[";
        for (int64_t r = 0; r < rows; ++r) {
            if (r) s.push_back(',');
            json_number(cols[c][c == 0 ? r : r * stride], s);
        }
        s.push_back(']');
    }
    s.push_back('}');
    *len = (int64_t)s.size();
    if (buf && cap > (int64_t)s.size()) std::memcpy(buf, s.c_str(), s.size() + 1);
    return MR_OK;
    MR_GUARD_END
}

int mr_ray_result_append_tsv(const char *path, const double *t, const double *x,
                             const double *y, const double *kx, const double *ky,
                             int64_t n_rows, int64_t stride)
{
    MR_GUARD_BEGIN
    if (!path || n_rows < 0 || stride < 1 || (n_rows > 0 && (!t || !x || !y || !kx || !ky)))
        return mr::set_error(MR_ERR_INVALID_ARGUMENT, "invalid arrays");
    std::string s = "t x y kx ky\n";
    for (int64_t r = 0; r < n_rows; ++r) {
        const int64_t o = r * stride;
        if (std::isnan(x[o])) break;                  // ray.rs:231-233: x only
        display_number(t[r], s);
        s.push_back(' ');
        const double row[4] = {x[o], y[o], kx[o], ky[o]};
        for (double v : row) {
            display_number(v, s);
            s.push_back(' ');
        }
        s += " \n";                                   // writeln!(" ")
    }
    s += "END\n";
    FILE *f = std::fopen(path, "ab");                 // create or append
    if (!f) return mr::set_error(MR_ERR_IO, std::string("cannot open ") + path);
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    if (std::fclose(f) != 0 || !ok)
        return mr::set_error(MR_ERR_IO, std::string("cannot write ") + path);
    return MR_OK;
    MR_GUARD_END
}

}  // extern "C"
