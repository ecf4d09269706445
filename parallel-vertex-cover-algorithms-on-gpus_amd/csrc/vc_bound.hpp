// vc_bound.hpp -- the edge bound of a search-tree node, shared by the gfx950
// kernel and host code (no HIP types, so a plain C++ compiler can include it).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define VC_HD __host__ __device__
#else
#define VC_HD
#endif

namespace vcg {

// Prune test of a node after the rule fixpoint (src/GlobalWorkList.cuh:132-139
// for MVC, src/GlobalWorkListParameterized.cuh:135-137 for PVC).
//
//   edges  live edges of the node
//   d      vertices already in the partial cover
//   T      the threshold the high-degree rule ran with: every vertex with
//          deg + d >= T went into the cover, so afterwards every live
//          degree is at most T - d - 1
//   B      the search looks for covers smaller than B: the live incumbent
//          (MVC) or k + 1 (PVC)
//
// Below this node a cover smaller than B adds at most B - d - 1 vertices,
// each covering at most T - d - 1 live edges, so more edges than
// (B - d - 1)(T - d - 1) rule it out.  The reference prunes at
// edges >= (B - d - 1)^2 + 1, which is the same test when B == T (always for
// PVC, T = k + 1).  For MVC it reads the incumbent again after the rules
// (src/GlobalWorkList.cuh:101 and :135): when another worker lowered it in
// between (B < T), degrees up to T - d - 1 survive the rules and the square
// can prune a node whose subtree holds a smaller cover (a K_{2,3} remnant with
// T = d + 5 and B = d + 3: 6 edges >= 2^2 + 1, yet its two hubs give d + 2).
// Taking the product with the threshold the rules used keeps the bound sound.
// B above T (impossible while the incumbent only falls) is clamped to T.
VC_HD inline bool bound_prunes(uint32_t edges, uint32_t d, uint32_t T, uint32_t B)
{
    if (B > T) B = T;
    if (d >= B) return true;
    return (uint64_t)edges > (uint64_t)(B - d - 1u) * (uint64_t)(T - d - 1u);
}

}  // namespace vcg
