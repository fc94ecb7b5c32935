/*
 * spobs.h -- C ABI of the observable-file writer: correlation measurements in
 * the reference's HDF5 .obs layout, so that tools reading SpinParser output
 * (its Python reader, test evaluators) read this engine's output unchanged.
 *
 * Layout (src/SU2/SU2MeasurementCorrelation.cpp:179-355,
 * src/TRI/TRIMeasurementCorrelation.cpp:318-493):
 *   /<group>/meta/latticeVectors   n_bravais x float[3]   (HDF5 array type)
 *   /<group>/meta/basis            B x float[3]
 *   /<group>/meta/sites            B x Rb x float[3]      (positions of the sites in range)
 *   /<group>/data/measurement_<k>  group with float attribute "cutoff" (1 element)
 *                                  and float dataset "data" of shape (B, Rb)
 * <k> counts the measurement groups already present; a measurement at a
 * cutoff already in the file is discarded (the reference logs a warning).
 * Host-only, no GPU; links the HDF5 C library.
 */
#ifndef SPOBS_H
#define SPOBS_H

#ifdef __cplusplus
extern "C" {
#endif

#define SP_OBS_OK 0
#define SP_OBS_DUPLICATE 1   /* cutoff already present: nothing written          */
#define SP_OBS_EARG -1       /* invalid argument                                  */
#define SP_OBS_EIO -2        /* HDF5 open / create / write failure                */

/* Append one correlation measurement to group `group` of the file `path`
 * (created if absent), writing the meta group first when it does not exist;
 * replaces SU2MeasurementCorrelation::_writeOutfileCorrelation and
 * _writeOutfileHeader (src/SU2/SU2MeasurementCorrelation.cpp:179-355). */
int sp_obs_write_correlation(const char *path, const char *group, float cutoff, int n_bravais,
                             const float *bravais, int B, const float *basis, int Rb, const float *sites,
                             const float *data);

/* Number of measurement groups in /<group>/data (0 if absent), and reading one
 * back (cutoff attribute and B*Rb data values); for tests and tools. */
int sp_obs_count(const char *path, const char *group);
int sp_obs_read_correlation(const char *path, const char *group, int k, float *cutoff, float *data, int n);

/* Checkpoint files in the reference's data-file layout
 * (src/SU2/SU2EffectiveAction.hpp:79-205, src/TRI/TRIEffectiveAction.hpp:79-205):
 *   /checkpoint_<k>  group with float attribute "cutoff" and float datasets
 *                    "cutoff" (1), "v2" (N_w) and the vertex arrays ("v4dd",
 *                    "v4ss" for SU(2); "v4" for TRI)
 * sp_ckpt_write appends (append != 0 and the file exists) or truncates;
 * returns the checkpoint id written, SP_OBS_DUPLICATE_CKPT (-3) if a
 * checkpoint at this cutoff exists, or an error code (< 0).  Values are stored
 * as float, as in the reference.  sp_ckpt_read reads checkpoint `id` (id < 0:
 * the last one) into the arrays, each of the given size. */
#define SP_OBS_DUPLICATE_CKPT -3
int sp_ckpt_write(const char *path, int append, float cutoff, int narrays, const char *const *names,
                  const long long *sizes, const float *const *data);
int sp_ckpt_read(const char *path, int id, float *cutoff, int narrays, const char *const *names,
                 const long long *sizes, float *const *data);
/* Number of checkpoints in a data file (0 if the file is absent); the
 * post-processing stage reads checkpoints 0, 1, ... until readCheckpoint
 * reports false (src/SpinParser.cpp:192-197). */
int sp_ckpt_count(const char *path);

/* Message of the last failure in this thread. */
const char *sp_obs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SPOBS_H */
