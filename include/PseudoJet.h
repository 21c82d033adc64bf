/*
 * PseudoJet.h — the particle/jet record of the drop-in interface.
 *
 * Same layout as the reference's record (PseudoJet.h:4-11): int index, bool isJet,
 * then px, py, pz, E as doubles; 40 bytes with 3 padding bytes after isJet.  The
 * layout is what callers of cluster() put in device memory, so it is kept as is.
 */
#ifndef PseudoJet_h
#define PseudoJet_h

struct PseudoJet {
  int index;
  bool isJet;
  double px;
  double py;
  double pz;
  double E;
};

static_assert(sizeof(PseudoJet) == 40, "PseudoJet must keep the 40-byte layout of the reference");

#endif  // PseudoJet_h
