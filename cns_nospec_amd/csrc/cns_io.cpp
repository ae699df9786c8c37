// cns_io.cpp -- the reference program's text formats (include/cns_io.h).
//
// Host-only code.  Field files of big boxes are large (a 512^3 box is 703 M values,
// about 17 GB of text), so parsing and formatting are split over threads; the
// bytes written are exactly what the reference's fprintf("%.17e\t") loops write.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>

#include "cns_host.h"
#include "cns_io.h"

namespace {

using cns::fail;

struct Padded {
    size_t xg, pg, cg;
};

bool padded_of(const int dim_g[3], Padded &p)
{
    if (!dim_g || dim_g[0] <= 0 || dim_g[1] <= 0 || dim_g[2] <= 0) return false;
    p.xg = (size_t(dim_g[0]) + 15) / 16 * 16;
    p.pg = p.xg * size_t(dim_g[1]);
    p.cg = p.pg * size_t(dim_g[2]);
    return true;
}

unsigned worker_count(size_t work_units)
{
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    if (hw > 16) hw = 16;
    if (work_units < hw) hw = unsigned(work_units ? work_units : 1);
    return hw;
}

template <class F>
void parallel_for(unsigned nthreads, F &&body)
{
    if (nthreads <= 1) {
        body(0u);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (unsigned t = 0; t < nthreads; t++) th.emplace_back([&body, t] { body(t); });
    for (auto &x : th) x.join();
}

// Whole file in memory, NUL-terminated so strtod/strtol stop at the end.
int slurp(const char *path, std::vector<char> &buf)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(CNS_EIO, "cannot open %s: %s", path, std::strerror(errno));
    buf.clear();
    if (std::fseek(f, 0, SEEK_END) == 0) {
        long sz = std::ftell(f);
        if (sz >= 0) buf.reserve(size_t(sz) + 1);
        std::rewind(f);
    }
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    bool err = std::ferror(f);
    std::fclose(f);
    if (err) return fail(CNS_EIO, "read error on %s", path);
    buf.push_back('\0');
    return CNS_OK;
}

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Next integer token at *pos (after whitespace).
bool next_int(const char *&pos, long &v)
{
    while (is_ws(*pos)) pos++;
    char *e;
    errno = 0;
    v = std::strtol(pos, &e, 10);
    if (e == pos || errno || !(is_ws(*e) || *e == '\0')) return false;
    pos = e;
    return true;
}

bool next_double(const char *&pos, double &v)
{
    while (is_ws(*pos)) pos++;
    char *e;
    v = std::strtod(pos, &e);
    if (e == pos || !(is_ws(*e) || *e == '\0')) return false;
    pos = e;
    return true;
}

}  // namespace

extern "C" {

size_t cns_field_len(const int dim_g[3], int ncomp)
{
    Padded p;
    if (ncomp <= 0 || !padded_of(dim_g, p)) return 0;
    return size_t(ncomp) * p.cg;
}

int cns_read_general_input(const char *path, cns_config *cfg)
{
    if (!path || !cfg) return fail(CNS_EINVAL, "NULL argument");
    std::vector<char> buf;
    if (int rc = slurp(path, buf)) return rc;
    const char *pos = buf.data();
    cns_config c{};
    long iv[9];
    const char *names[] = {"ng", "nc", "ncells", "lo[0]", "lo[1]", "lo[2]", "hi[0]", "hi[1]", "hi[2]"};
    for (int i = 0; i < 9; i++)
        if (!next_int(pos, iv[i])) return fail(CNS_EIO, "%s: cannot parse %s", path, names[i]);
    c.ng = int(iv[0]);
    c.nc = int(iv[1]);
    c.ncells = int(iv[2]);
    for (int d = 0; d < 3; d++) {
        c.lo[d] = int(iv[3 + d]);
        c.hi[d] = int(iv[6 + d]);
    }
    for (int d = 0; d < 3; d++)
        if (!next_double(pos, c.dx[d])) return fail(CNS_EIO, "%s: cannot parse dx[%d]", path, d);
    if (!next_double(pos, c.cfl)) return fail(CNS_EIO, "%s: cannot parse cfl", path);
    if (!next_double(pos, c.eta)) return fail(CNS_EIO, "%s: cannot parse eta", path);
    if (!next_double(pos, c.alam)) return fail(CNS_EIO, "%s: cannot parse alam", path);
    long ns;
    if (!next_int(pos, ns)) return fail(CNS_EIO, "%s: cannot parse nsteps", path);
    c.nsteps = int(ns);
    if (!next_double(pos, c.dt)) return fail(CNS_EIO, "%s: cannot parse dt", path);
    if (c.ng < 0 || c.nc <= 0) return fail(CNS_EIO, "%s: ng = %d, nc = %d", path, c.ng, c.nc);
    for (int d = 0; d < 3; d++) {
        c.dim[d] = c.hi[d] - c.lo[d] + 1;
        c.dim_g[d] = c.dim[d] + 2 * c.ng;
        if (c.dim[d] <= 0) return fail(CNS_EIO, "%s: hi[%d] < lo[%d]", path, d, d);
    }
    *cfg = c;
    return CNS_OK;
}

int cns_write_general_input(const char *path, const cns_config *cfg)
{
    if (!path || !cfg) return fail(CNS_EINVAL, "NULL argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(CNS_EIO, "cannot open %s: %s", path, std::strerror(errno));
    const cns_config &c = *cfg;
    std::fprintf(f, "%d\n%d\n%d\n", c.ng, c.nc, c.ncells);
    std::fprintf(f, "%d %d %d\n", c.lo[0], c.lo[1], c.lo[2]);
    std::fprintf(f, "%d %d %d\n", c.hi[0], c.hi[1], c.hi[2]);
    std::fprintf(f, "%.17e %.17e %.17e\n", c.dx[0], c.dx[1], c.dx[2]);
    std::fprintf(f, "%.17e\n%.17e\n%.17e\n", c.cfl, c.eta, c.alam);
    std::fprintf(f, "%d\n%.17e\n", c.nsteps, c.dt);
    bool err = std::ferror(f);
    if (std::fclose(f) != 0 || err) return fail(CNS_EIO, "write error on %s", path);
    return CNS_OK;
}

int cns_read_field(const char *path, const int dim_g[3], int ncomp, int header, double *host)
{
    Padded P;
    if (!path || !host || ncomp <= 0 || !padded_of(dim_g, P))
        return fail(CNS_EINVAL, "bad argument (path, host, ncomp or dim_g)");
    std::vector<char> buf;
    if (int rc = slurp(path, buf)) return rc;
    const char *pos = buf.data();
    if (header) {
        long h[4];
        for (int i = 0; i < 4; i++)
            if (!next_int(pos, h[i])) return fail(CNS_EIO, "%s: malformed header", path);
        if (h[0] != ncomp || h[1] != dim_g[0] || h[2] != dim_g[1] || h[3] != dim_g[2])
            return fail(CNS_EIO, "%s: header says %ld components of %ld x %ld x %ld, expected %d of %d x %d x %d",
                        path, h[0], h[1], h[2], h[3], ncomp, dim_g[0], dim_g[1], dim_g[2]);
    }
    const size_t d0 = size_t(dim_g[0]), d1 = size_t(dim_g[1]), d2 = size_t(dim_g[2]);
    const size_t expect = size_t(ncomp) * d0 * d1 * d2;
    const char *begin = pos, *end = buf.data() + buf.size() - 1;  // end points at the NUL
    const size_t span = size_t(end - begin);
    const unsigned T = worker_count(span >> 22);  // about 4 MB of text per worker
    // Chunk boundaries moved forward onto whitespace so that no token straddles two chunks.
    std::vector<const char *> cut(T + 1);
    cut[0] = begin;
    cut[T] = end;
    for (unsigned t = 1; t < T; t++) {
        const char *c = begin + span / T * t;
        if (c < cut[t - 1]) c = cut[t - 1];
        while (c < end && !is_ws(*c)) c++;
        cut[t] = c;
    }
    std::vector<std::vector<double>> vals(T);
    std::vector<const char *> bad(T, nullptr);
    parallel_for(T, [&](unsigned t) {
        const char *p = cut[t], *e = cut[t + 1];
        auto &v = vals[t];
        v.reserve(size_t(e - p) / 24 + 1);
        while (true) {
            while (p < e && is_ws(*p)) p++;
            if (p >= e) break;
            char *q;
            double x = std::strtod(p, &q);
            if (q == p || !(is_ws(*q) || *q == '\0')) {
                bad[t] = p;
                break;
            }
            v.push_back(x);
            p = q;
        }
    });
    size_t total = 0;
    for (unsigned t = 0; t < T; t++) {
        if (bad[t]) {
            char tok[32];
            std::snprintf(tok, sizeof tok, "%.24s", bad[t]);
            return fail(CNS_EIO, "%s: value %zu is not a number (\"%s\")", path, total + vals[t].size(), tok);
        }
        total += vals[t].size();
    }
    if (total != expect)
        return fail(CNS_EIO, "%s: expected %zu values (%d x %zu x %zu x %zu), found %zu", path, expect, ncomp,
                    d2, d1, d0, total);
    std::vector<size_t> first(T + 1, 0);
    for (unsigned t = 0; t < T; t++) first[t + 1] = first[t] + vals[t].size();
    parallel_for(T, [&](unsigned t) {
        size_t n = first[t];
        for (double x : vals[t]) {
            size_t i = n % d0, r = n / d0;
            size_t j = r % d1;
            r /= d1;
            size_t k = r % d2, l = r / d2;
            host[l * P.cg + k * P.pg + j * P.xg + i] = x;
            n++;
        }
    });
    return CNS_OK;
}

int cns_write_field(const char *path, const int dim_g[3], int ncomp, int header, const double *host)
{
    Padded P;
    if (!path || !host || ncomp <= 0 || !padded_of(dim_g, P))
        return fail(CNS_EINVAL, "bad argument (path, host, ncomp or dim_g)");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(CNS_EIO, "cannot open %s: %s", path, std::strerror(errno));
    if (header) {
        std::fprintf(f, "%d\n", ncomp);
        std::fprintf(f, "%d %d %d\n", dim_g[0], dim_g[1], dim_g[2]);
    }
    const int d0 = dim_g[0], d1 = dim_g[1], d2 = dim_g[2];
    // Planes (l, k) are formatted in parallel batches and written in order: each plane is
    // d1 rows of d0 "%.17e\t" values and "\n", then "\n"; each component ends with "\n".
    const size_t nplanes = size_t(ncomp) * size_t(d2);
    const unsigned T = worker_count(size_t(d0) * d1 * nplanes >> 16);
    const size_t batch = size_t(T) * 4;
    std::vector<std::string> text(batch);
    bool err = false;
    for (size_t p0 = 0; p0 < nplanes && !err; p0 += batch) {
        const size_t np = std::min(batch, nplanes - p0);
        parallel_for(T, [&](unsigned t) {
            for (size_t b = t; b < np; b += T) {
                const size_t pl = p0 + b, l = pl / size_t(d2), k = pl % size_t(d2);
                std::string &s = text[b];
                s.clear();
                s.reserve(size_t(d0 + 1) * d1 * 25 + 2);
                char num[40];
                const double *plane = host + l * P.cg + k * P.pg;
                for (int j = 0; j < d1; j++) {
                    const double *row = plane + size_t(j) * P.xg;
                    for (int i = 0; i < d0; i++) {
                        int len = std::snprintf(num, sizeof num, "%.17e\t", row[i]);
                        s.append(num, size_t(len));
                    }
                    s.push_back('\n');
                }
                s.push_back('\n');
                if (k == size_t(d2) - 1) s.push_back('\n');
            }
        });
        for (size_t b = 0; b < np && !err; b++)
            err = std::fwrite(text[b].data(), 1, text[b].size(), f) != text[b].size();
    }
    err = err || std::ferror(f);
    if (std::fclose(f) != 0 || err) return fail(CNS_EIO, "write error on %s", path);
    return CNS_OK;
}

}  // extern "C"

namespace {

int make_dir(const std::string &d)
{
    if (mkdir(d.c_str(), 0755) == 0 || errno == EEXIST) return CNS_OK;
    return fail(CNS_EIO, "cannot create directory %s: %s", d.c_str(), std::strerror(errno));
}

}  // namespace

extern "C" {

int cns_write_plotfile(const char *dir, int ncomp, int nbox, const int *lo, const int *hi, int ng,
                       const double *host, const double prob_lo[3], const double prob_hi[3],
                       const double dx[3], double time)
{
    if (!dir || !*dir || ncomp <= 0 || ncomp > 99 || nbox <= 0 || !lo || !hi || ng < 0 || !host ||
        !prob_lo || !prob_hi || !dx)
        return fail(CNS_EINVAL, "bad argument (dir, ncomp, nbox, boxes, ng, host or geometry)");
    int dlo[3], dhi[3];
    std::vector<size_t> offs(size_t(nbox) + 1, 0);  // each box's start in host
    for (int b = 0; b < nbox; b++) {
        int dg[3];
        for (int d = 0; d < 3; d++) {
            if (hi[3 * b + d] < lo[3 * b + d]) return fail(CNS_EINVAL, "box %d is empty along %d", b, d);
            dg[d] = hi[3 * b + d] - lo[3 * b + d] + 1 + 2 * ng;
            dlo[d] = b ? std::min(dlo[d], lo[3 * b + d]) : lo[3 * b + d];
            dhi[d] = b ? std::max(dhi[d], hi[3 * b + d]) : hi[3 * b + d];
        }
        offs[b + 1] = offs[b] + cns_field_len(dg, ncomp);
    }
    const std::string root(dir), lev = root + "/Level_00";
    int rc = make_dir(root);
    if (rc == CNS_OK) rc = make_dir(lev);
    if (rc != CNS_OK) return rc;

    // Cell_D_00000: every box's FAB; the valid cells only, component by component
    std::vector<long long> fab_at(static_cast<size_t>(nbox));
    std::vector<double> mn(size_t(nbox) * ncomp), mx(size_t(nbox) * ncomp);
    {
        const std::string path = lev + "/Cell_D_00000";
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) return fail(CNS_EIO, "cannot open %s: %s", path.c_str(), std::strerror(errno));
        bool err = false;
        std::vector<double> buf;
        for (int b = 0; b < nbox && !err; b++) {
            const int *l = lo + 3 * b, *h = hi + 3 * b;
            const int n0 = h[0] - l[0] + 1, n1 = h[1] - l[1] + 1, n2 = h[2] - l[2] + 1;
            const int dg[3] = {n0 + 2 * ng, n1 + 2 * ng, n2 + 2 * ng};
            Padded P;
            padded_of(dg, P);
            fab_at[b] = std::ftell(f);
            err = std::fprintf(f,
                               "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
                               "((%d,%d,%d) (%d,%d,%d) (0,0,0)) %d\n",
                               l[0], l[1], l[2], h[0], h[1], h[2], ncomp) < 0;
            buf.resize(size_t(n0) * n1 * n2);
            for (int c = 0; c < ncomp && !err; c++) {
                const double *src = host + offs[b] + size_t(c) * P.cg;
                size_t at = 0;
                double lo_v = HUGE_VAL, hi_v = -HUGE_VAL;
                for (int k = 0; k < n2; k++)
                    for (int j = 0; j < n1; j++) {
                        const double *row = src + size_t(k + ng) * P.pg + size_t(j + ng) * P.xg + ng;
                        for (int i = 0; i < n0; i++) {
                            const double v = row[i];
                            buf[at++] = v;
                            lo_v = std::min(lo_v, v);
                            hi_v = std::max(hi_v, v);
                        }
                    }
                mn[size_t(b) * ncomp + c] = lo_v;
                mx[size_t(b) * ncomp + c] = hi_v;
                // x86-64 and the GPU hosts are little-endian, as the FAB header declares
                err = std::fwrite(buf.data(), sizeof(double), buf.size(), f) != buf.size();
            }
        }
        err = err || std::ferror(f);
        if (std::fclose(f) != 0 || err) return fail(CNS_EIO, "write error on %s", path.c_str());
    }
    // Cell_H: the FabArray header
    {
        const std::string path = lev + "/Cell_H";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return fail(CNS_EIO, "cannot open %s: %s", path.c_str(), std::strerror(errno));
        std::fprintf(f, "1\n1\n%d\n0\n(%d 0\n", ncomp, nbox);
        for (int b = 0; b < nbox; b++)
            std::fprintf(f, "((%d,%d,%d) (%d,%d,%d) (0,0,0))\n", lo[3 * b], lo[3 * b + 1], lo[3 * b + 2],
                         hi[3 * b], hi[3 * b + 1], hi[3 * b + 2]);
        std::fprintf(f, ")\n%d\n", nbox);
        for (int b = 0; b < nbox; b++) std::fprintf(f, "FabOnDisk: Cell_D_00000 %lld\n", fab_at[b]);
        for (const std::vector<double> *v : {&mn, &mx}) {
            std::fprintf(f, "\n%d,%d\n", nbox, ncomp);
            for (int b = 0; b < nbox; b++) {
                for (int c = 0; c < ncomp; c++) std::fprintf(f, "%.17e,", (*v)[size_t(b) * ncomp + c]);
                std::fprintf(f, "\n");
            }
        }
        const bool err = std::ferror(f);
        if (std::fclose(f) != 0 || err) return fail(CNS_EIO, "write error on %s", path.c_str());
    }
    // Header: the plotfile's own header
    {
        const std::string path = root + "/Header";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) return fail(CNS_EIO, "cannot open %s: %s", path.c_str(), std::strerror(errno));
        std::fprintf(f, "NavierStokes-V1.1\n%d\n", ncomp);
        for (int c = 0; c < ncomp; c++) std::fprintf(f, "Variable_%02d\n", c + 1);
        std::fprintf(f, "3\n%.17e\n0\n", time);
        std::fprintf(f, "%.17e %.17e %.17e\n", prob_lo[0], prob_lo[1], prob_lo[2]);
        std::fprintf(f, "%.17e %.17e %.17e\n", prob_hi[0], prob_hi[1], prob_hi[2]);
        std::fprintf(f, "\n((%d,%d,%d) (%d,%d,%d) (0,0,0))\n0\n", dlo[0], dlo[1], dlo[2], dhi[0], dhi[1], dhi[2]);
        std::fprintf(f, "%.17e %.17e %.17e\n0\n0\n", dx[0], dx[1], dx[2]);
        std::fprintf(f, "0 %d %.17e\n0\n", nbox, time);
        for (int b = 0; b < nbox; b++)
            for (int d = 0; d < 3; d++)
                std::fprintf(f, "%.17e %.17e\n", prob_lo[d] + dx[d] * lo[3 * b + d],
                             prob_lo[d] + dx[d] * (hi[3 * b + d] + 1));
        std::fprintf(f, "Level_00/Cell\n");
        const bool err = std::ferror(f);
        if (std::fclose(f) != 0 || err) return fail(CNS_EIO, "write error on %s", path.c_str());
    }
    return CNS_OK;
}

}  // extern "C"
