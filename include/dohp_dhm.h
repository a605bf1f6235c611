/*
 * dohp_dhm.h -- DHM (HDF5) field files of block function spaces: the dohp viewer's on-disk format
 * (src/viewer/dhm.c, src/fs/impls/cont/contviewdhm.c) for the spaces and closure vectors of this library.
 *
 * A separate library (dohp_amd/libdohp_dhm.so, host code linked with HDF5), so that the operator library does not
 * depend on HDF5.  Every entry point returns a dErr (0 = success; dohpDHMGetLastError gives the text) and takes host
 * pointers: copy device vectors to the host first.
 *
 * Layout (as the reference's): groups /dohp/{mesh,fs,step,type}; committed types string, Units_type, BoundingBox_type,
 * Time_type, Field_type, FS_type, Vec_type in /dohp/type with the reference's member names; a function space is a
 * one-record FS_type dataset /dohp/fs/<name> referring to its mesh; a time step is a group /dohp/step/<%03d> with a
 * "time" attribute (Time_type); a vector is a dataset [n/bs][bs] of doubles in its step with a "meta" attribute
 * (Vec_type: region reference to its space, time, units, internal state).  The mesh record of a block space is the
 * string "block M0 M1 M2" (the reference stores the name of the MOAB file) beside a coordinates dataset [nvert][3],
 * and the FS_type strings (tag names of a MOAB mesh in the reference) hold the block space's degree, numbering and
 * Dirichlet sets.
 */
#ifndef DOHP_DHM_H
#define DOHP_DHM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dErr;
typedef struct _p_dohpDHM *dohpDHM;

/* PetscFileMode of the viewer (dhm.c:264-283). */
enum { dohpDHM_READ = 0, dohpDHM_WRITE = 1, dohpDHM_APPEND = 2 };

void dohpDHMGetLastError(char *buf, size_t len);

/* Opens a DHM file (PetscViewerCreate_DHM + PetscViewerFileSetMode + PetscViewerFileSetName + dViewerDHMSetUp,
 * src/viewer/dhm.c:439-482,368-375,348-357,251-302): WRITE truncates and creates the /dohp groups, READ and APPEND open them. */
dErr dohpDHMOpen(const char *path, int mode, dohpDHM *v);
dErr dohpDHMClose(dohpDHM *v);
/* Time of the next step written (dViewerDHMSetTime, dhm.c:377-390): each call starts a new step. */
dErr dohpDHMSetTime(dohpDHM v, double time);
/* Units of the step times (dViewerDHMSetTimeUnits, dhm.c:392-402). */
dErr dohpDHMSetTimeUnits(dohpDHM v, const char *units, double scale);
/* Number of steps and their times, at most cap of them (dViewerDHMGetSteps, dhm.c:527-553). */
dErr dohpDHMGetSteps(dohpDHM v, int *nsteps, double *times_host, int cap);
/* Step that later reads look in (dViewerDHMSetTimeStep, dhm.c:484-499). */
dErr dohpDHMSetTimeStep(dohpDHM v, int step);

/* Writes the record of a block function space (dFSView_Cont_DHM, contviewdhm.c:58-111): an M0 x M1 x M2 block of
 * degree `degree`, Dirichlet sets `dirichlet_mask` (bit 1 walls, 2 top, 4 bottom), vertex coordinates
 * [(M0+1)(M1+1)(M2+1)][3], bs fields named fieldnames[bs] (NULL: "u0", "u1", ...).  A second write of the same
 * name is a no-op (the reference writes one record per space). */
dErr dohpDHMWriteBlockSpace(dohpDHM v, const char *fsname, int M0, int M1, int M2, int degree, int dirichlet_mask,
                            const double *vcoords_host, int bs, const char *const *fieldnames);
/* Reads it back (dFSLoadIntoFS_Cont_DHM, contviewdhm.c:228-348): M[3], degree, mask, bs; vcoords_host (may be
 * NULL) receives the vertex coordinates. */
dErr dohpDHMReadBlockSpace(dohpDHM v, const char *fsname, int *M, int *degree, int *dirichlet_mask, int *bs,
                           double *vcoords_host);
/* Writes a closure vector [nclosure] (owned nodes first, then Dirichlet; bs interleaved) of space fsname into the
 * current step (VecView_Dohp_FSCont_DHM, contviewdhm.c:132-206); unit_scale is its units' scale. */
dErr dohpDHMWriteVec(dohpDHM v, const char *fsname, const char *vecname, long nclosure, int bs, const double *x_host,
                     double unit_scale);
/* Reads the vector of that name from the current step into x_host [nclosure] (VecDohpLoadIntoVector,
 * contviewdhm.c:350-401); the stored size must match. */
dErr dohpDHMReadVec(dohpDHM v, const char *vecname, long nclosure, int bs, double *x_host);

#ifdef __cplusplus
}
#endif
#endif
