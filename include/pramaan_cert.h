/* pramaan_cert.h -- C entry points of the certificate writer and checker
 * (libpramaan_cxx.so), for hosts that do not link the C++ interface.
 *
 * They replace, for a pm_lp in memory, the reference's
 *   generate_certificate + write_certificate
 *       (src/certificate/certificate.cpp:23-153), and
 *   pramaan-verify <model.mps> <certificate.cert>
 *       (src/certificate/verifier_main.cpp:28-229).
 * V1 files are the reference's PRAMAAN_CERTIFICATE_V1 text, byte for byte;
 * V2 files add the FP64 dual y and its residuals (pramaan/certificate.hpp). */
#ifndef PRAMAAN_CERT_H
#define PRAMAAN_CERT_H

#include "pramaan_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Writes a certificate for x (length lp->n).  y == NULL writes V1 (the dual
 * fields UNAVAILABLE); otherwise y (length lp->m, reference sign: reduced
 * cost c + A^T y) is written and V2's dual residual, complementarity
 * residual and objective bound gap are recomputed from (x, y) in FP64.
 * Returns 0, or 1 with the reason in errbuf. */
int pm_cert_write(const pm_lp* lp, const double* x, const double* y, const char* status, double tolerance,
                  const char* path, char* errbuf, int errlen);

/* Reads the certificate at path and checks it against lp: 0 = valid,
 * 1 = invalid (each failed check on its own line in msg), 2 = unreadable
 * (the parse error in msg). */
int pm_cert_verify(const pm_lp* lp, const char* path, char* msg, int msglen);

#ifdef __cplusplus
}
#endif

#endif /* PRAMAAN_CERT_H */
