/* qsv_sdk.h -- QRunes programs on the MI355X engine: the C functions of
 * QPanda's SDK facade (QPandaSDK/QPandaAPI/QPandaAPI.h:136-173) plus handle
 * based equivalents.  Library: qpanda_amd/libqpanda_sdk.so (links libqsv.so).
 *
 * Return codes are QPanda's QError values (QuantumInstructionHandle/QError.h):
 * 2 = qErrorNone, 4 = qParameterError, 5 = qbitError, 9 = setComputeUnitError,
 * 7 = initStateError, 10 = runProgramError, 13 = notReadFileError.
 *
 * String out-parameters follow the reference: with *len == 0 only the length
 * is returned; otherwise up to *len bytes are copied (no terminating NUL) and
 * *len is set to the full length. */
#ifndef QSV_SDK_H
#define QSV_SDK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- the reference's process-wide SDK instance (QPandaAPI.cpp:308-372) ---- */

/* QPandaSDK::loadFile (QPandaAPI.cpp:58-95): parse a QRunes file. */
int loadFileAPI(char *pcFilePath);
/* QPandaSDK::setComputeUnit (QPandaAPI.cpp:97-121): 2 = this engine. */
int setComputeUnitAPI(int iComputeUnit);
/* QPandaSDK::run (QPandaAPI.cpp:123-205): iRepeat runs from |0...0>. */
int runAPI(int iRepeat);
/* file name of the loaded program */
int getFileNameAPI(char *buf, int *pLength);
/* QPandaSDK::getResult (QPandaAPI.cpp:207-229): "<key> <value>\n" lines */
int getResultAPI(char *buf, int *pLength);
/* QPandaSDK::getQuantumState (QPandaAPI.cpp:231-246): state text of the last run */
int getQuantumStateAPI(char *buf, int *pLength);

/* ---- the front end alone (QRunesParserAPI.cpp:28-44) ---- */

/* Parse a QRunes file or text.  Returns the QINIT qubit count, or -1 on any
 * error.  buf receives one line per instruction of the flattened list
 * ("<kind> q... | angle=<%.17g> dagger=<0|1> ctrl=<a,b> creg=<c>"). */
int qrunes_parse_file(const char *path, char *buf, int *len);
int qrunes_parse_text(const char *text, char *buf, int *len);

/* ---- handle-based SDK instances (no reference counterpart) ---- */

typedef struct qpanda_sdk qpanda_sdk;
qpanda_sdk *qpanda_sdk_create(void);
void qpanda_sdk_destroy(qpanda_sdk *h);
int qpanda_sdk_load_file(qpanda_sdk *h, const char *path);
int qpanda_sdk_set_compute_unit(qpanda_sdk *h, int unit);
/* seed the compute unit's Park-Miller stream (after set_compute_unit) */
void qpanda_sdk_set_seed(qpanda_sdk *h, int32_t seed);
/* 1: always re-run the program per shot, as the reference does */
void qpanda_sdk_set_resample_disabled(qpanda_sdk *h, int off);
int qpanda_sdk_run(qpanda_sdk *h, int repeat);
/* 1 when the last run sampled its shots from one execution */
int qpanda_sdk_last_run_sampled(qpanda_sdk *h);
int qpanda_sdk_get_result(qpanda_sdk *h, char *buf, int *len);
int qpanda_sdk_get_state(qpanda_sdk *h, char *buf, int *len);
/* The reference CLI's state files of the last run's final state
 * (QCommand.cpp:402-487): binary 0 writes path + ".fullstate" (the state text
 * as getQuantumState prints it), 1 writes path + ".fullstate.binary" (2^n
 * records {f64 re, f64 im}).  Returns a QError code. */
int qpanda_sdk_save_state(qpanda_sdk *h, const char *path, int binary);
/* FP64 amplitudes (re, im pairs) of the last run's final state: copies up to
 * count amplitudes into out and returns the state's length (0 before a run). */
int64_t qpanda_sdk_get_amplitudes(qpanda_sdk *h, double *out, int64_t count);

#ifdef __cplusplus
}
#endif

#endif
