/* psm_host.h -- C ABI of libpsm_host.so, the host side of the poststackmig
 * driver: configuration file, SEG-Y ingest and output, velocity providers and
 * the whole command-line run.  The migration itself goes through libpsm_hip.so
 * (psm.h); nothing here migrates on the CPU.
 *
 * Status codes are those of psm.h (PSM_OK, PSM_ERR_*); a failing call writes a
 * message into the caller's `err` buffer (may be NULL) when it takes one, or
 * keeps it for psmh_last_error() otherwise (per thread).
 */
#ifndef PSM_HOST_H
#define PSM_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "psm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing psmh_* call on this thread ("" if none). */
const char* psmh_last_error(void);

/* ------------------------------------------------------------ configuration
 * Replaces ConfigParser::parseConfigFile + validateConfig
 * (config_parser.cpp:74-212, fields config_parser.h:7-20): `key=value` lines,
 * '#' comments, surrounding quotes stripped; required input_data, output_data,
 * inline_step, crossline_step, velocity; defaults angle_aperture 30,
 * amp_correction true, paddings 0, n_threads 0.  Same validation messages. */
typedef struct psmh_config {
  char input_data[1024];
  char output_data[1024];
  char velocity[1024];
  double inline_step, crossline_step;
  int inline_padding, crossline_padding;
  double angle_aperture;
  int amp_correction;
  int n_threads;
} psmh_config;

int psmh_config_parse(const char* path, psmh_config* out);

/* ------------------------------------------------------------------ SEG-Y in
 * Replaces SegyFile's metadata read and lookup table
 * (segy_utils.cpp:39-111, 152-214): 3200-byte text header, 400-byte binary
 * header (dt in microseconds @16, samples @20, format @24: 1 = IBM float,
 * 5 = IEEE float, anything else read as 5), fixed-length traces, inline /
 * crossline labels from trace-header bytes 188 / 192 (big-endian int32);
 * traces with a zero label are skipped and a later trace with the same
 * labels replaces an earlier one.  Labels are sorted ascending. */
typedef struct psmh_segy psmh_segy;

int psmh_segy_open(const char* path, psmh_segy** out);
void psmh_segy_close(psmh_segy* s);
/* n_traces = traces in the file; n_il / n_xl = distinct labels; fmt = 1 or 5 */
int psmh_segy_info(const psmh_segy* s, size_t* n_traces, size_t* n_t, int* dt_us, int* fmt,
                   size_t* n_il, size_t* n_xl);
const int32_t* psmh_segy_inline_labels(const psmh_segy* s);
const int32_t* psmh_segy_crossline_labels(const psmh_segy* s);
/* Whole-volume ingest: every trace decoded once into cube[n_il][n_xl][n_t]
 * (float32, missing traces zero).  Replaces the per-window re-reads of
 * readIlineByCoords / readTraceData (segy_utils.cpp:272-383,
 * migration_kernel.cpp:441-467). */
int psmh_segy_read_cube(psmh_segy* s, float* cube);

/* IBM single <-> IEEE float (segy_utils.cpp:243-270): the decoder is exact,
 * the encoder truncates the fraction as the reference does. */
void psmh_ibm_to_ieee(const uint32_t* ibm, float* out, size_t n);
void psmh_ieee_to_ibm(const float* x, uint32_t* out, size_t n);

/* ----------------------------------------------------------------- SEG-Y out
 * Replaces SegyFile::createOutputFile + SegyOutputFile::writeTrace
 * (segy_utils.cpp:398-487, 598-648) and the write loop of main.cpp:385-415:
 * text header copied; binary header copied with samples @20 and format @24
 * set; traces in output order (inline-major); a trace whose output inline is
 * a real inline of `in` and whose (inline, crossline) exists in `in` carries
 * that input trace's 240-byte header, any other trace a zero header with only
 * the labels at 188 / 192; samples in the input's format, big-endian. */
int psmh_segy_write(const char* path, const psmh_segy* in, const int32_t* out_il_labels,
                    size_t n_out_il, const int32_t* out_xl_labels, size_t n_out_xl,
                    const float* image /* [n_out_il][n_out_xl][n_t] */, size_t n_t);

/* -------------------------------------------------------------- velocities
 * Replaces VelocityProviderFactory::create and the providers'
 * getVelocityIline as prepareVelocitySliceWithPadding uses them
 * (velocity_reader.cpp:12-35, 203-261, 264-421; migration_kernel.cpp:470-511):
 * the Vrms cube vel[n_il][n_xl][n_t] (host memory) on the data grid given by
 * the label lists and (n_t, dt).  `param` is a number (constant velocity), a
 * path containing ".sgy"/".segy" (any case: SEG-Y velocity, exact traces
 * copied, otherwise linear in time and an inverse-(1 + |dil| + |dxl|) mean of
 * the 2 x 2 neighbouring traces), or a text table "INLINE CROSSLINE TIME VEL"
 * (inverse-(1 + d) mean over all points, evaluated on the GPU of `ctx`). */
int psmh_velocity_cube(psm_ctx* ctx, const char* param, const int32_t* il_labels, size_t n_il,
                       const int32_t* xl_labels, size_t n_xl, size_t n_t, double dt,
                       float* vel);

/* --------------------------------------------------------------- whole run
 * The `poststackmig <config>` command (main.cpp:66-443): parse and validate
 * the configuration, ingest the input cube once, build the Vrms cube, migrate
 * the whole padded volume on GPU `device` (psm_migrate_volume), write the
 * output SEG-Y.  Progress goes to stdout, errors to stderr as
 * "Error: <message>".  Returns the process exit status (0 or 1). */
int psmh_run(const char* config_path, int device);

#ifdef __cplusplus
}
#endif

#endif /* PSM_HOST_H */
