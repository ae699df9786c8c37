/*
 * cns_io.h -- the reference program's text formats, for a drop-in driver
 * (libcns_hip.so; host-side functions, no device work).
 *
 * general_input: whitespace-separated, in this order (cpu/helper_functions.cpp:42-52,
 * gpu/util_cpu.cpp:315-325): ng, nc, ncells, lo[3], hi[3], dx[3], cfl, eta, alam,
 * nsteps, dt.  lo/hi are the file's 0-based interior bounds; the interior extent is
 * dim = hi - lo + 1 and the ghosted extent dim_g = dim + 2 ng.
 *
 * Field files (multistep_input, output): nc components of the ghosted box, component
 * slowest, then k, j, and i fastest, as "%.17e\t" per value with a newline after every
 * row, after every plane and after every component (print_4D, cpu/helper_functions.cpp:85-98,
 * gpu/util_cpu.cpp:171-184; read_3D, gpu/util_cpu.cpp:76-85).  "%.17e" round-trips every
 * double exactly.  The output file starts with a header line "nc" and a line
 * "dim_g[0] dim_g[1] dim_g[2]" (cpu/main.cpp:36-40, gpu/main.cu:55-58); multistep_input
 * has none.
 *
 * Host buffers are in the ABI's padded layout of a ghosted array (cns.h):
 * idx = l*cg + k*pg + j*xg + i, xg = PAD16(dim_g[0]), so one copy moves them to or from
 * the device.  Padding lanes are neither read nor written.
 */
#ifndef CNS_IO_H
#define CNS_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CNS_EIO (-6) /* a file could not be opened, read or written, or is malformed */

typedef struct cns_config {
    int ng, nc, ncells;
    int lo[3], hi[3]; /* as written in the file */
    double dx[3];
    double cfl, eta, alam;
    int nsteps;
    double dt;
    int dim[3], dim_g[3]; /* derived by cns_read_general_input */
} cns_config;

/* Parses a general_input file into *cfg and derives dim and dim_g.
 * Replaces: read_configurations(h_const)  cpu/helper_functions.cpp:38-83;
 *           read_configurations(h_const, argc, argv)  gpu/util_cpu.cpp:287-364 (file part). */
int cns_read_general_input(const char *path, cns_config *cfg);

/* Writes *cfg as a general_input file that cns_read_general_input and the
 * reference's read_configurations both parse (dx, cfl, eta, alam and dt as "%.17e").
 * No counterpart in the reference, whose inputs were written by external tools. */
int cns_write_general_input(const char *path, const cns_config *cfg);

/* Reads ncomp components of a ghosted field of extents dim_g into host (padded layout).
 * header = 1 expects and checks the output file's two header lines (nc == ncomp and
 * the extents); header = 0 reads a bare body (multistep_input).  Exactly
 * ncomp*dim_g[0]*dim_g[1]*dim_g[2] values must follow; anything short or unparsable
 * is CNS_EIO.  Parsing is split over threads.
 * Replaces: the read_3D loop over components  cpu/main.cpp:22-25, gpu/main.cu:39-42
 *           (read_3D: cpu/helper_functions.cpp:166-174, gpu/util_cpu.cpp:76-85). */
int cns_read_field(const char *path, const int dim_g[3], int ncomp, int header, double *host);

/* Writes ncomp components of a ghosted field from host (padded layout) in the print_4D
 * format, byte-identical to the reference's fprintf output, with the output header when
 * header = 1.  Formatting is split over threads.
 * Replaces: the output block  cpu/main.cpp:36-40, gpu/main.cu:55-58 (print_4D:
 *           cpu/helper_functions.cpp:85-98, gpu/util_cpu.cpp:171-184). */
int cns_write_field(const char *path, const int dim_g[3], int ncomp, int header,
                    const double *host);

/* Number of doubles of a padded ghosted host buffer of ncomp components (what
 * cns_read_field and cns_write_field index), or 0 for invalid extents.
 * No counterpart in the reference. */
size_t cns_field_len(const int dim_g[3], int ncomp);

/* Writes the directory `dir` as a BoxLib plotfile of one level: the valid (interior)
 * cells of nbox boxes, ncomp components named Variable_01, Variable_02, ...
 * Box b covers the 0-based cells lo[3b..3b+2] .. hi[3b..3b+2] of the domain (their
 * bounding box is the problem domain); `host` holds the boxes back to back, box b as a
 * padded ghosted array of its extents with ng ghost cells (the layout of include/cns.h;
 * nbox = 1 is a single box, a cns_domain array is nbox equal boxes).  prob_lo/prob_hi
 * are the physical corners of the domain, dx the cell size, time the solution time.
 * Files: dir/Header (format "NavierStokes-V1.1": names, dimension, time, finest level 0,
 * prob_lo, prob_hi, problem domain, dx, each box's physical extent, "Level_00/Cell"),
 * dir/Level_00/Cell_H (the box array, the FAB offsets, per-box per-component min and
 * max) and dir/Level_00/Cell_D_00000 (per box a "FAB ((8, (64 11 52 0 1 12 0 1023)),
 * (8, (8 7 6 5 4 3 2 1)))((lo) (hi) (0,0,0)) ncomp" line, then its ncomp components of
 * little-endian doubles, x fastest).  dir is created if absent.
 * Replaces: write_plotfile  fortran90/write_plotfile.f90:11-53 (U at step istep into
 * "plt%05d", fortran90/main.f90:122-125, :163-168), whose file writer is BoxLib's
 * fabio_ml_multifab_write_d (not in the reference tree: the layout above follows the
 * published BoxLib plotfile format and is not checked byte for byte). */
int cns_write_plotfile(const char *dir, int ncomp, int nbox, const int *lo, const int *hi, int ng,
                       const double *host, const double prob_lo[3], const double prob_hi[3],
                       const double dx[3], double time);

#ifdef __cplusplus
}
#endif

#endif /* CNS_IO_H */
