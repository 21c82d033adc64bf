/*
 * cluster.h — the drop-in engine interface (replaces the reference's cluster.h:4-6).
 *
 * void cluster(PseudoJet *particles, int size, Algorithm algo, double r);
 *   particles: DEVICE pointer to `size` PseudoJet records (caller allocates/frees).
 *   On return, in place: records with isJet == true hold the inclusive jets
 *   (no pT cut); the jet promoted while n pseudojets were live sits in slot n-1;
 *   index == slot for every record.  Other records have isJet == false.
 *   The call is ordered on the null stream: a following hipMemcpy sees the result.
 *   Errors print "file, line: call; name: message" to stderr and abort(), like the
 *   reference's cudaCheck (cudaCheck.h:11-47).
 * Algorithm's value is the generalised-kt exponent p (cluster.h:4 of the reference).
 * Which engine runs is chosen at link time, as in the reference: the bucket table's
 * choice per event (tri_matrix binary, lib/libfjmi_cluster.so), dense (n_array),
 * tiled (grid), or the per-step launch floor with no result (tri_matrix_empty).
 */
#ifndef cluster_h
#define cluster_h

#include "PseudoJet.h"

enum class Algorithm { Kt = 1, CambridgeAachen = 0, AntiKt = -1 };

void cluster(PseudoJet *particles, int size, Algorithm algo, double r);

#endif  // cluster_h
