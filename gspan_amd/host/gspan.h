// gspan.h -- mining entry point with the reference's signature (gspan.h:41 in
// haozixu/gSpan): mines every frequent connected pattern of g with at most k edges under
// MNI support and prints one line per pattern, "(from to fl el tl)...: support"
// (gspan.cu:67-77), in the reference's order.  The work runs on the GPU through
// include/gspan_amd.h.
#ifndef GSPAN_AMD_GSPAN_H
#define GSPAN_AMD_GSPAN_H

#include "graph.h"

void fsm(const CSRGraph& g, int k, int min_support);

// Same, with the execution mode (GS_MODE_*) and MNI semantics (GS_MNI_*) of gspan_amd.h.
// Throws std::runtime_error with the library's message on failure.
void fsm(const CSRGraph& g, int k, int min_support, int mode, int mni_mode);

#endif
