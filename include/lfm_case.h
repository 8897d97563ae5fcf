/*
 * lfm_case.h -- C ABI of the OpenFOAM case input/output around the MI355X
 * time step: the polyMesh (ASCII or binary format), the dictionaries
 * CInputReader reads and the
 * initial fields, as the reference takes them from OpenFOAM through
 * CPolyMeshReaderOF / CDictReaderOF / CRunTimeManagerOF
 * (polyMeshReaderOF/polyMeshReaderOF.cpp, dictReaderOF/dictReaderOF.cpp,
 * runTimeManagerOF/runTimeManagerOF.cpp, src/inputReader.cpp:7-64), and the
 * time-directory writer (Mesh::solve's runTime.write(), src/mesh_solver.cpp:
 * 821-845).  Host code only; no device is touched.
 *
 * Same conventions as lfm_hip.h: 0 or a negative LFM_E* code, message in
 * lfm_last_error().  Arrays handed out by a case stay valid until
 * lfm_case_close.
 */
#ifndef LFM_CASE_H
#define LFM_CASE_H

#include <stdint.h>

#include "lfm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lfm_case lfm_case;

/* Open an OpenFOAM case directory.  rank < 0 reads the undecomposed case
 * (constant/polyMesh, <start time>/); rank >= 0 reads processor<rank>/ of a
 * decomposed case (its polyMesh with processor patches, faceProcAddressing,
 * optional cellSubmesh, and its <start time>/ fields), with the dictionaries
 * of the case root.  The start time follows controlDict startFrom as
 * Foam::Time reads it: startTime (the entry; the default), firstTime or
 * latestTime (the smallest / largest numeric time directory of the case or
 * processor directory, a restart).  Replaces CRunTimeManagerOF(rank) +
 * CPolyMeshReaderOF + CInputReader::initialize (info/lfm_solve.cpp:39-114,
 * runTimeManagerOF/runTimeManagerOF.cpp:5-12). */
int lfm_case_open(const char *case_dir, int32_t rank, lfm_case **out);
int lfm_case_close(lfm_case *c);

/* The mesh (patch roles resolved with lfm_patch_role, processor patches
 * with their neighbour rank) and, for a processor case, the partition
 * (face keys from |faceProcAddressing| - 1). part may be NULL. */
int lfm_case_mesh(lfm_case *c, lfm_polymesh *mesh, lfm_partition *part);
/* CInputReader's keys and defaults (src/inputReader.cpp:13-64) plus
 * controlDict deltaT / adjustTimeStep; lfm/doublePrecision false selects
 * LFM_PREC_FP32P. */
int lfm_case_params(lfm_case *c, lfm_params *prm);
/* Initial p, T, U (n_cells*3) and alpha (NULL when the case has none) of the
 * start time (cfd_v0.cpp:919-935). */
int lfm_case_fields(lfm_case *c, const double **p, const double **T, const double **U3, const double **alpha);
/* The start time (as startFrom selects it), controlDict endTime, deltaT,
 * writeInterval (time steps when
 * writeControl is timeStep, else seconds) and writeControl (0 timeStep,
 * 1 runTime/adjustableRunTime). */
int lfm_case_time(lfm_case *c, double *start, double *end, double *delta_t, double *write_interval,
                  int32_t *write_control);
/* Any scalar of a case dictionary, as CDictReaderOF::readDictScalar:
 * dict is "<dir>/<name>" relative to the case root (e.g. "system/controlDict"),
 * subdict a '/'-separated path ("" for the top level); when the key is
 * missing and allow_default, *out = dflt, else LFM_EINVAL. */
int lfm_case_dict_scalar(lfm_case *c, const char *dict, const char *subdict, const char *key, double dflt,
                         int32_t allow_default, double *out);

/* Write p, U and T of one time into <case>[/processor<rank>]/<time name>/
 * (ASCII volScalarField / volVectorField, internalField nonuniform, the
 * boundaryField entries of the start-time files, T from p / (rho R)), with
 * rho and E, and the sponge distance alpha when the case has one (the
 * reference reads it AUTO_WRITE, polyMeshReaderOF.cpp:700-709, so a
 * restart from this time finds it).
 * q is SoA [(D+2)][n_cells] conservatives in original cell order. */
int lfm_case_write_time(lfm_case *c, double time, const double *q, int32_t dim);

/* The post-processing controls of controlDict lfm/post (haveAverage,
 * tStartAverage, haveForces, saveForcesStep; src/inputReader.cpp:19-23) as an
 * lfm_post for the start time: average_start = (int)((tStartAverage -
 * startTime) / deltaT) as Mesh::solve computes it (mesh_solver.cpp:253-254),
 * forces_own_face = 0. */
int lfm_case_post(lfm_case *c, lfm_post *post);

/* Write one scalar field (n_cells values, original cell order) into the time
 * directory, boundaries "calculated" (createScalarField fields such as pAvg
 * and pRMS, mesh_solver.cpp:266-271). */
int lfm_case_write_scalar(lfm_case *c, double time, const char *name, const char *dimensions, const double *v);

#ifdef __cplusplus
}
#endif
#endif
