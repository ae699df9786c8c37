// cns_run -- the reference program's main loop on libcns_hip.so.
//
//   cns_run [-d testcases_dir] [-o output] [-n nsteps] [-q] [-p plot_int [-P plot_dir]] [name]
//
// Reads <dir>/general_input and <dir>/multistep_input (or <dir>/<name>_general_input
// and <dir>/<name>_multistep_input when a name is given; dir defaults to
// ../testcases), copies U to the device, advances nsteps RK3 steps with cns_advance's
// kernels, prints "dt, courno = %le, %le" per step (as the reference's new_advance)
// and "Total time: %lf", and writes the ghosted U to `output` in the reference's
// format.  Follows gpu/main.cu:11-68 and cpu/main.cpp:9-58; the timed region, as in
// gpu/main.cu:44-52, spans the copy in, the steps and the copy out.
//
// With -p plot_int > 0 it also writes BoxLib plotfiles of U as the Fortran driver
// does (fortran90/main.f90:122-125, :163-168): plot_dir/plt00000 before the first
// step, then plt%05d after every step that is a multiple of plot_int and after the
// last, with time = the sum of the steps' dt and the domain centred on the origin
// (prob_lo = -n dx / 2, prob_hi = n dx / 2: -0.1 and 0.1 for the reference's inputs,
// fortran90/main.f90:93-106).  Writing them is not part of the timed region.
//
// Exit codes: 0 done, 1 usage / input / output error, 2 library or HIP error,
// 3 the state became non-physical (non-finite Courant number).
#include <hip/hip_runtime.h>
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cns.h"
#include "cns_io.h"

namespace {

double now_s()
{
    timeval tv;
    gettimeofday(&tv, nullptr);
    return double(tv.tv_sec) + 1e-6 * double(tv.tv_usec);
}

int usage(const char *argv0)
{
    std::fprintf(stderr, "usage: %s [-d testcases_dir] [-o output] [-n nsteps] [-q] [-p plot_int [-P plot_dir]] [name]\n",
                 argv0);
    return 1;
}

struct DeviceBuffers {
    double *p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double *scalars = nullptr;
    ~DeviceBuffers()
    {
        for (double *q : p)
            if (q) (void)hipFree(q);
        if (scalars) (void)hipFree(scalars);
    }
};

#define HIP_OR_DIE(call)                                                                           \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "cns_run: %s: %s\n", #call, hipGetErrorString(e_));               \
            return 2;                                                                              \
        }                                                                                          \
    } while (0)

}  // namespace

int main(int argc, char **argv)
{
    std::string dir = "../testcases", out = "output", name, plot_dir = ".";
    int nsteps_override = -1, plot_int = 0;
    bool quiet = false;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        if ((s == "-d" || s == "-o" || s == "-n" || s == "-p" || s == "-P") && a + 1 < argc) {
            std::string v = argv[++a];
            if (s == "-d") dir = v;
            else if (s == "-o") out = v;
            else if (s == "-P") plot_dir = v;
            else if (s == "-p") plot_int = std::atoi(v.c_str());
            else nsteps_override = std::atoi(v.c_str());
        } else if (s == "-q") {
            quiet = true;
        } else if (s == "-h" || s == "--help" || (!s.empty() && s[0] == '-')) {
            return usage(argv[0]);
        } else if (name.empty()) {
            name = s;
        } else {
            return usage(argv[0]);
        }
    }
    const std::string prefix = dir + "/" + (name.empty() ? "" : name + "_");
    const std::string cfg_path = prefix + "general_input", in_path = prefix + "multistep_input";

    cns_config cfg;
    if (cns_read_general_input(cfg_path.c_str(), &cfg) != CNS_OK) {
        std::fprintf(stderr, "cns_run: %s\n", cns_last_error_string());
        return usage(argv[0]);
    }
    if (nsteps_override >= 0) cfg.nsteps = nsteps_override;
    if (cfg.ng != 4 || cfg.nc != 5) {
        std::fprintf(stderr, "cns_run: %s: ng = %d, nc = %d; this build supports ng = 4, nc = 5\n",
                     cfg_path.c_str(), cfg.ng, cfg.nc);
        return 1;
    }
    const size_t lenU = cns_field_len(cfg.dim_g, 5), lenQ = cns_field_len(cfg.dim_g, 6);
    const size_t lenDF = size_t(5) * ((size_t(cfg.dim[0]) + 15) / 16 * 16) * size_t(cfg.dim[1]) * size_t(cfg.dim[2]);
    std::vector<double> U(lenU, 0.0);
    if (cns_read_field(in_path.c_str(), cfg.dim_g, 5, 0, U.data()) != CNS_OK) {
        std::fprintf(stderr, "cns_run: %s\n", cns_last_error_string());
        return 1;
    }

    DeviceBuffers dev;
    const size_t lens[5] = {lenU, lenU, lenQ, lenDF, lenDF};
    for (int b = 0; b < 5; b++) {
        HIP_OR_DIE(hipMalloc(&dev.p[b], lens[b] * sizeof(double)));
        HIP_OR_DIE(hipMemset(dev.p[b], 0, lens[b] * sizeof(double)));
    }
    HIP_OR_DIE(hipMalloc(&dev.scalars, 2 * sizeof(double)));
    cns_ctx *ctx = nullptr;
    if (cns_ctx_create(&ctx, cfg.dim, cfg.dx, cfg.cfl, cfg.eta, cfg.alam) != CNS_OK) {
        std::fprintf(stderr, "cns_run: %s\n", cns_last_error_string());
        return 2;
    }

    // plotfile of the device U at step istep (fortran90/write_plotfile.f90:11-53)
    double plot_s = 0.0, time = 0.0;
    auto plot = [&](int istep) -> int {
        const double t0 = now_s();
        if (hipMemcpy(U.data(), dev.p[0], lenU * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
            std::fprintf(stderr, "cns_run: copying U for plotfile %d failed\n", istep);
            return 2;
        }
        char nm[32];
        std::snprintf(nm, sizeof nm, "/plt%05d", istep);
        double plo[3], phi[3];
        for (int d = 0; d < 3; d++) {
            phi[d] = 0.5 * cfg.dim[d] * cfg.dx[d];
            plo[d] = -phi[d];
        }
        if (cns_write_plotfile((plot_dir + nm).c_str(), 5, 1, cfg.lo, cfg.hi, cfg.ng, U.data(), plo, phi,
                               cfg.dx, time) != CNS_OK) {
            std::fprintf(stderr, "cns_run: %s\n", cns_last_error_string());
            return 1;
        }
        plot_s += now_s() - t0;
        return 0;
    };

    int rc = 0;
    double total = -now_s();
    HIP_OR_DIE(hipMemcpy(dev.p[0], U.data(), lenU * sizeof(double), hipMemcpyHostToDevice));
    if (plot_int > 0) rc = plot(0);
    for (int s = 0; s < cfg.nsteps && rc == 0; s++) {
        if (cns_advance_async(ctx, dev.p[0], dev.p[1], dev.p[2], dev.p[3], dev.p[4], dev.scalars,
                              dev.scalars + 1) != CNS_OK) {
            std::fprintf(stderr, "cns_run: step %d: %s\n", s + 1, cns_last_error_string());
            rc = 2;
            break;
        }
        double h[2];
        HIP_OR_DIE(hipMemcpy(h, dev.scalars, sizeof h, hipMemcpyDeviceToHost));
        if (!quiet) std::printf("dt, courno = %le, %le\n", h[0], h[1]);
        if (!std::isfinite(h[1])) {
            std::fprintf(stderr, "cns_run: step %d: Courant number %g is not finite\n", s + 1, h[1]);
            rc = 3;
            break;
        }
        time += h[0];
        if (plot_int > 0 && ((s + 1) % plot_int == 0 || s + 1 == cfg.nsteps)) rc = plot(s + 1);
    }
    if (rc == 0) HIP_OR_DIE(hipMemcpy(U.data(), dev.p[0], lenU * sizeof(double), hipMemcpyDeviceToHost));
    total += now_s() - plot_s;
    cns_ctx_destroy(ctx);
    if (rc) return rc;
    std::printf("Total time: %lf\n", total);
    std::fflush(stdout);
    if (cns_write_field(out.c_str(), cfg.dim_g, 5, 1, U.data()) != CNS_OK) {
        std::fprintf(stderr, "cns_run: %s\n", cns_last_error_string());
        return 1;
    }
    return 0;
}
