/*
 * xdem_io.h — state files of the DEM step: the reference's result and resume
 * formats, and checkpoints that resume bit-exactly.
 *
 * Reference formats (little-endian, no padding):
 *   part%04d.bin   dem_simulation::saveResult / saveResult_f
 *                  (source/simulation/dem_simulation.cpp:491-547):
 *                  double ct, uint32 np, pos[np*4] (double; float for the
 *                  FP32 model — the file records no precision, the size tells)
 *   <name>_final.bfr  xDynamicsSolver::saveFinalResult (xDynamicsSolver.cpp:111-130)
 *                  + dem_simulation::saveFinalResult(_f) (:551-583):
 *                  double ct, int32 precision (2 = FP64, 1 = FP32), int32 flag = 1,
 *                  pos[np*4], vel[np*3], avel[np*3]; np is not stored (the
 *                  reader, startingModel::setDEMData, startingModel.cpp:32-60,
 *                  takes it from the model).
 * The reference resumes from the .bfr with zero accelerations and F = m g
 * (dem_simulation.cpp:240-251, 585-593), so its resume is not bit-identical to
 * an uninterrupted run.  The checkpoints below (no reference counterpart) add
 * what that needs: accelerations, force/moment, the step count, prescribed
 * motion counters and moved triangles (simulation); the scene table and the
 * tangential-spring history (batch).
 *
 * Checkpoint container: char magic[8] "XDEMCKP1", uint32 kind (1 simulation,
 * 2 batch), uint32 version (1), uint64 payload bytes, the payload, uint64
 * FNV-1a hash of the payload.  Readers reject a wrong magic, kind or version,
 * a size that does not match, and a hash mismatch (truncated or altered files).
 */
#ifndef XDEM_IO_H
#define XDEM_IO_H

#include "xdem.h"
#include "xdem_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- part%04d.bin (host arrays) */
int32_t xdem_write_part_f64(const char* path, double ct, uint32_t np, const double* pos4);
int32_t xdem_write_part_f32(const char* path, double ct, uint32_t np, const float* pos4);
/* Reads either precision (from the size); pos4 may be NULL to query ct/np;
 * cap = entries pos4 can take; *precision = 2 (FP64) or 1 (FP32, widened). */
int32_t xdem_read_part(const char* path, double* ct, uint32_t* np, int32_t* precision,
                       double* pos4, uint32_t cap);

/* ---- <name>_final.bfr (host arrays; the DEM block only) */
int32_t xdem_write_bfr_f64(const char* path, double ct, uint32_t np, const double* pos4,
                           const double* vel3, const double* avel3);
int32_t xdem_write_bfr_f32(const char* path, double ct, uint32_t np, const float* pos4,
                           const float* vel3, const float* avel3);
/* np from the model, as the reference reader; FP32 files are widened. */
int32_t xdem_read_bfr(const char* path, uint32_t np, double* ct, int32_t* precision,
                      double* pos4, double* vel3, double* avel3);

/* ---- the FP64 simulation object (xdem_sim) */
/* dem_simulation::saveResult: <dir>/part%04d.bin of the current positions,
 * ct = steps * dt. */
int32_t xdem_sim_save_part(const xdem_sim* s, const char* dir, uint32_t part);
/* saveFinalResult: the .bfr of the current state. */
int32_t xdem_sim_save_final(const xdem_sim* s, const char* path);
/* setStartingData after initialize: positions, velocities and angular
 * velocities from a .bfr; accelerations zero, F = m g, T = 0; the step
 * count becomes round(ct / dt). */
int32_t xdem_sim_load_bfr(xdem_sim* s, const char* path);
/* Checkpoint of the whole dynamic state.  Loading needs a simulation built
 * the same way (particles, walls, triangles and motions in the same order);
 * stepping on from it is bit-identical to never having stopped. */
int32_t xdem_sim_save_checkpoint(const xdem_sim* s, const char* path);
int32_t xdem_sim_load_checkpoint(xdem_sim* s, const char* path);

/* ---- the batched FP32c path (xdem_batch) */
/* Every scene: its description (planes included), state and tangential
 * history. */
int32_t xdem_batch_save_checkpoint(xdem_batch* b, const char* path);
/* Adds the checkpoint's scenes to b (in checkpoint order); their new ids go to
 * scene_ids (cap entries), their number to *nscenes.  Stepping on is
 * bit-identical to never having stopped (same block lengths). */
int32_t xdem_batch_load_checkpoint(xdem_batch* b, const char* path, uint32_t* scene_ids,
                                   uint32_t cap, uint32_t* nscenes);

/* Container check without loading: *kind = 1 (simulation) or 2 (batch). */
int32_t xdem_checkpoint_verify(const char* path, uint32_t* kind);

#ifdef __cplusplus
}  // extern "C"
#endif
#endif
