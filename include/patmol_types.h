/*
 * patmol_types.h -- data layouts shared with Patmol's C driver.
 *
 * These structs keep the reference's field order and sizes so that a
 * reference driver, its basis model and this library can pass them to each
 * other unchanged:
 *   Shell  -- shell.h:9-13   (npr, lsh, cen, off, nst, al*, co*)
 *   Center -- center.h:5-7   (q, r[3]), positions in bohr
 *   ERIFS  -- erifs.h:13-17  (rank, nproc, nfiles, numints*, basename*)
 *   ERINT  -- erifs.h:19-22  (16-byte record: double val; u16 i, j, k, l)
 * When a translation unit already includes the reference's own shell.h /
 * center.h / erifs.h, define PATMOL_HAVE_REFERENCE_TYPES before including
 * this header to use those definitions instead.
 */
#ifndef PATMOL_TYPES_H
#define PATMOL_TYPES_H

#ifndef PATMOL_HAVE_REFERENCE_TYPES
typedef struct {
    int npr, lsh, cen, off, nst;
    double *al;
    double *co;
} Shell;

typedef struct {
    double q, r[3];
} Center;

typedef struct {
    int rank, nproc, nfiles;
    int *numints;
    char *basename;
} ERIFS;

typedef struct {
    double val;
    unsigned short i, j, k, l;
} __attribute__((aligned(16))) ERINT;
#endif

#endif
