// swg_state.h -- device-resident state of an ensemble of planetary systems.
//
// Layout in HBM: every per-body quantity is a separate array (SoA) so that a
// wavefront touching 64 consecutive bodies issues fully coalesced loads.
// Systems are contiguous segments: massive bodies of system s occupy
// [pl_off[s], pl_off[s+1]) and test particles [tp_off[s], tp_off[s+1]).  A
// single large system is an ensemble of one; every kernel is written for the
// segmented case.
//
// Encounter lists are the reference's (index1, index2) lists
// (encounter_module.f90:23-47) concatenated system by system, each ordered
// index1 ascending then index2 ascending (encounter_check.f90:436-475), with
// global 0-based body indices.  They are produced from a bit matrix: word
// (row i, block b) holds the 64 encounter flags of pairs (i, 64*b .. 64*b+63).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swg {

struct Vec3p {
    double *x, *y, *z;
};

// Massive bodies ("pl") of all systems.
struct PlDev {
    Vec3p rh, vh, vb, rbeg, rend;
    double *Gm, *R, *rhill, *renc;
    int32_t *status, *levelg, *levelm, *sys;
    int32_t *tiny;  // 1: Gmass < GMTINY (SyMBA), interacts only with the fully interacting bodies; may be NULL
};

// Test particles ("tp") of all systems.
struct TpDev {
    Vec3p rh, vh, vb;
    int32_t *status, *levelg, *levelm, *sys;
};

// Per-system scalars.
struct SysDev {
    int32_t *pl_off, *tp_off;     // nsys+1 each
    double *GMcb;                 // nsys
    double *cbvb, *ptbeg, *ptend; // 3*nsys, component-major per system (s*3+d)
    int32_t *lfirst;              // 1 = first kick (vh->vb) pending
    int32_t *enc_pp, *enc_pt;     // level-0 encounter flags of the current step
    int32_t *flag_pp, *flag_pt;   // per-substep list-check flags
    int32_t *err;                 // error bits: 1 collision, 2 pl drift failure
    int32_t *frozen;              // 1 = system left out of the steps (its state is kept as is)
    int32_t *nplm;                // fully interacting massive bodies (the first nplm of the system); may be NULL
};

// A (pl, pl) or (pl, tp) encounter list with its bit matrix and adjacency.
struct ListDev {
    // bit matrix
    uint64_t *words;   // per-system blocks, row-major: words[woff[s] + (i - off_s) * nb_s + b]
    int32_t *wpre;     // exclusive popcount prefix of each word within its row
    int64_t *woff;     // nsys
    int32_t *nb;       // words per row, per system
    int32_t *rowcnt;   // pairs per row (index1)
    int64_t *rowstart; // exclusive scan of rowcnt
    // pair list (length n)
    int32_t *i1, *i2;
    int8_t *lvdotr, *status, *level;
    // adjacency of the index2 side: pairs whose index2 is body b, in list order
    int32_t *colcnt;
    int64_t *colstart;
    int32_t *col;      // pair ids
};

}  // namespace swg
