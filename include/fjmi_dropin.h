/*
 * fjmi_dropin.h — additions beside the drop-in cluster() (include/cluster.h),
 * served by the same library context.
 *
 * int inclusive_jets(const PseudoJet *particles, int size, double ptmin,
 *                    PseudoJet *jets, int *count);
 *   After cluster(): the records of `particles` (device memory) with isJet and
 *   px^2 + py^2 >= ptmin^2, whole and sorted by decreasing px^2 + py^2 (equal
 *   values keep slot order), are written to `jets` (device memory, room for size
 *   records) and their number to *count (device memory).  Ordered on the null
 *   stream.  The reference's driver does this filter and sort on the host after
 *   copying every record back (main.cc:309-327).  Errors abort like cluster().
 */
#ifndef fjmi_dropin_h
#define fjmi_dropin_h

#include "PseudoJet.h"

void inclusive_jets(const PseudoJet *particles, int size, double ptmin, PseudoJet *jets, int *count);

#endif  // fjmi_dropin_h
