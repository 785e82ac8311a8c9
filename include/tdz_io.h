/* Restart files and observables traces in HDF5 (libtdz_io.so, host only).
 *
 * Separate from libtdz.so so that the propagator does not depend on the HDF5
 * library; this one links libhdf5 (1.10). All arrays are host arrays; complex
 * values are interleaved (re, im) doubles, vectors in the natural (x fastest)
 * ordering. Functions return 0 on success and an error code otherwise, the
 * message in tdz_io_last_error(); nothing throws across the ABI.
 */
#ifndef TDZ_IO_H
#define TDZ_IO_H

#ifdef __cplusplus
extern "C" {
#endif

const char *tdz_io_last_error(void);

/* Checkpoint (replaces TDSEZCheckpointer::Write / Read, src/checkpoint.cpp:101-216):
 * root attributes version (= 1), step, time, time_step, dofs, propagator
 * (0 cn, 1 pade22, 2 magnus4) and the dataset "psi" of shape [dofs][2]. Pade/Magnus
 * runs may add their stage guess history (CheckpointGuess, src/propagator.cpp:1546-1590):
 * attribute guess_count_<s> and datasets guess_<s>_<j> ([dofs][2]) for nslots slots,
 * guess[s * order + j] pointing at dofs complex values. The file is written to
 * <path>.tmp and renamed over <path>, so a crash never leaves a torn checkpoint. */
int tdz_io_checkpoint_write(const char *path, long long step, double time, double dt, int propagator,
                            long long dofs, const double *psi, int nslots, int order, const int *guess_count,
                            const double *const *guess);
/* Reads a checkpoint written for the same grid, time step and propagator (each is
 * checked, with the reference's messages); guess history is read when guess is not
 * null (guess_count[s] = 0 for slots the file does not hold). */
int tdz_io_checkpoint_read(const char *path, long long dofs, double dt, int propagator, long long *step,
                           double *time, double *psi, int nslots, int order, int *guess_count, double *const *guess);

/* Observables file (replaces the per-run HDF5 writer, src/manager.cpp:51-201):
 * extendable chunked datasets of real columns, one row per recorded step,
 *   dipoles          [t, E_a (dim), q<D_a> (dim), acceleration_a (dim)]  (1 + 3 dim)
 *   energies         [t, Kin, Pot, Int, Tot, <1/m>, Norm]               (7)
 *   autocorrelation  [t, Re<psi0|psi>, Im<psi0|psi>]                     (3)
 * (column layouts of src/monitor.cpp:1281-1330). open creates the file, or with
 * append = 1 extends an existing one (a restarted run). */
int tdz_io_trace_open(const char *path, int dim, int append, void **handle);
int tdz_io_trace_append(void *handle, long long rows, const double *dipoles, const double *energies,
                        const double *autocorr);
int tdz_io_trace_close(void *handle);
/* Reads a whole dataset: with out = null, only its shape. */
int tdz_io_trace_read(const char *path, const char *dataset, long long *rows, int *cols, double *out);

#ifdef __cplusplus
}
#endif

#endif /* TDZ_IO_H */
