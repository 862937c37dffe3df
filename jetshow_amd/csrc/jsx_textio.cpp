/*
 * jsx_textio.cpp -- the reference's text formats, host only:
 *
 *   jsx_write_map        write_2dvector (src/utils.cpp:211-224) into a file
 *                        opened for appending (src/main.cpp:1447-1462):
 *                        value/scale in the stream's default format (%g,
 *                        6 significant digits) and a space per value, a
 *                        newline per row -- the map_*.txt files the Python
 *                        drivers load (abc_simulations.py:188-263)
 *   jsx_write_stripe     write_vector (src/utils.cpp:227-236): one value and
 *                        a space per line (stripe_tau.txt,
 *                        example_simple.py:107,190-232)
 *   jsx_read_field_text  read_from_txt (src/utils.cpp:381-388): rows of
 *                        three numbers (z, r, value) until the first that
 *                        does not parse
 *   jsx_grid_from_points create_triangulation (src/SimulationInterpolater.cpp:
 *                        105-117) for points that form a complete regular
 *                        (z, r) grid: coordinates divided by pc, nodes laid
 *                        out for jsx_grid_upload
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/jetshow_amd.h"

namespace {

const double kPc = 3.0856775814671913e+18;   /* include/utils.h pc */

/* whitespace-separated numbers as the stream extraction operator reads them
 * (strtod accepts the same decimal and exponent forms, and inf / nan) */
struct Reader {
    FILE *f;
    std::string tok;
    static bool space(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
    bool next(double &v) {
        int c;
        do { c = fgetc(f); } while (c != EOF && space(c));
        if (c == EOF) return false;
        tok.clear();
        while (c != EOF && !space(c)) {
            tok.push_back((char)c);
            c = fgetc(f);
        }
        char *end = nullptr;
        v = strtod(tok.c_str(), &end);
        return end != tok.c_str() && *end == 0;
    }
};

/* sorted distinct values of a coordinate, equal within a relative 1e-12 of
 * the coordinate's range */
std::vector<double> distinct(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    std::vector<double> out;
    const double tol = 1e-12 * std::max(1.0, std::fabs(v.back() - v.front()));
    for (double x : v)
        if (out.empty() || x - out.back() > tol) out.push_back(x);
    return out;
}

int locate(const std::vector<double> &u, double x) {
    const double tol = 1e-12 * std::max(1.0, std::fabs(u.back() - u.front()));
    auto it = std::lower_bound(u.begin(), u.end(), x - tol);
    if (it == u.end() || std::fabs(*it - x) > tol) return -1;
    return (int)(it - u.begin());
}

/* uniform spacing: every step within 1e-9 of the mean step */
bool uniform(const std::vector<double> &u, double &d) {
    d = (u.back() - u.front()) / (double)(u.size() - 1);
    for (size_t i = 1; i < u.size(); ++i)
        if (std::fabs((u[i] - u[i - 1]) - d) > 1e-9 * std::fabs(d)) return false;
    return d > 0.0;
}

}  // namespace

extern "C" int jsx_write_map(const char *path, const double *plane, int64_t n_rows, int64_t n_cols, double scale,
                             int32_t append) {
    if (!path || n_rows < 0 || n_cols < 0 || (n_rows * n_cols > 0 && !plane)) return JSX_EINVAL;
    FILE *f = fopen(path, append ? "a" : "w");
    if (!f) return JSX_EINVAL;
    for (int64_t i = 0; i < n_rows; ++i) {
        for (int64_t j = 0; j < n_cols; ++j) fprintf(f, "%g ", plane[i * n_cols + j] / scale);
        fputc('\n', f);
    }
    return fclose(f) == 0 ? JSX_OK : JSX_EINVAL;
}

extern "C" int jsx_write_stripe(const char *path, const double *values, int64_t n, double scale, int32_t append) {
    if (!path || n < 0 || (n > 0 && !values)) return JSX_EINVAL;
    FILE *f = fopen(path, append ? "a" : "w");
    if (!f) return JSX_EINVAL;
    for (int64_t i = 0; i < n; ++i) fprintf(f, "%g \n", values[i] / scale);
    return fclose(f) == 0 ? JSX_OK : JSX_EINVAL;
}

extern "C" int jsx_read_field_text(const char *path, double *zrv, int64_t cap, int64_t *n_points) {
    if (!path || !n_points || cap < 0 || (cap > 0 && !zrv)) return JSX_EINVAL;
    FILE *f = fopen(path, "r");
    if (!f) return JSX_ENOTFOUND;
    Reader rd{f, {}};
    int64_t n = 0;
    double row[3];
    for (;;) {
        if (!rd.next(row[0]) || !rd.next(row[1]) || !rd.next(row[2])) break;
        if (n < cap)
            for (int k = 0; k < 3; ++k) zrv[3 * n + k] = row[k];
        ++n;
    }
    fclose(f);
    *n_points = n;
    return n > cap && cap > 0 ? JSX_EINVAL : JSX_OK;
}

extern "C" int jsx_grid_from_points(const double *zrv, int64_t n_points, double nan_value, jsx_grid_desc *desc,
                                    double *values, int64_t cap) {
    if (!zrv || !desc || n_points < 4) return JSX_EINVAL;
    std::vector<double> z(n_points), r(n_points);
    for (int64_t i = 0; i < n_points; ++i) {
        z[i] = zrv[3 * i] / kPc;
        r[i] = zrv[3 * i + 1] / kPc;
        if (!std::isfinite(z[i]) || !std::isfinite(r[i])) return JSX_EINVAL;
    }
    const std::vector<double> uz = distinct(z), ur = distinct(r);
    if (uz.size() < 2 || ur.size() < 2 || (int64_t)(uz.size() * ur.size()) != n_points) return JSX_EUNSUPPORTED;
    double dz, dr;
    if (!uniform(uz, dz) || !uniform(ur, dr)) return JSX_EUNSUPPORTED;
    const int64_t nz = (int64_t)uz.size(), nr = (int64_t)ur.size();
    if (nz > 0x7fffffff || nr > 0x7fffffff) return JSX_EINVAL;
    desc->nz = (int32_t)nz;
    desc->nr = (int32_t)nr;
    desc->z0 = uz.front();
    desc->dz = dz;
    desc->r0 = ur.front();
    desc->dr = dr;
    desc->nan_value = nan_value;
    if (!values) return JSX_OK;   /* shape query */
    if (cap < nz * nr) return JSX_EINVAL;
    std::vector<char> seen((size_t)(nz * nr), 0);
    for (int64_t i = 0; i < n_points; ++i) {
        const int a = locate(uz, z[i]), b = locate(ur, r[i]);
        if (a < 0 || b < 0) return JSX_EUNSUPPORTED;
        const size_t k = (size_t)a * (size_t)nr + (size_t)b;
        if (seen[k]) return JSX_EUNSUPPORTED;   /* a node twice: not a grid */
        seen[k] = 1;
        values[k] = zrv[3 * i + 2];
    }
    return JSX_OK;
}
