// Binary DRAT output of the simplifier (ParaFROST src/cpu/proof.cpp:92-232, the
// reference's -proof with its default binary format).
//
// The reference writes its proof lines one after another as its loops run; here
// the passes run in parallel, so each pass logs events tagged with the elected
// variable q that made them and a per-variable sequence number seq (the
// position in that variable's part of the sequential loop). Ordered by (q, seq)
// the events are the sequential order; decided and encoded after the pass
// (proof.hip), they give the reference's bytes.
//
// Event kinds:
//   PE_ADD / PE_DEL   'a' / 'd' line of the clause at pool reference `payload`
//                     (its literals at the end of the pass are those it had when
//                     the event was made: a clause is changed at most once per pass)
//   PE_UNIT           'a' line of the unit literal `payload`
//   PE_CAND_ADD       'a' line unless a literal is true when the line comes
//                     (addResolvent, proof.cpp:208-222) -- for passes whose units
//                     are only assigned after the pass (VE)
//   PE_CAND_UNIT      enqueueUnit if the literal is unassigned when it comes, a
//                     contradiction if it is false (bve.cpp:199-211)
//   PE_CONFLICT       a contradiction: the empty clause ends the proof
#pragma once
#include "device.hpp"

namespace pfg {

enum { PE_ADD = 1, PE_DEL = 2, PE_UNIT = 3, PE_CAND_ADD = 4, PE_CAND_UNIT = 5, PE_CONFLICT = 6 };

struct PEv {
    uint32_t q, seq, tag, pad;
    uint64_t payload;
};

struct ProofLog {
    PEv *ev;                   // null: proof output off
    unsigned long long *n;     // events logged (may exceed cap: checked by the host)
    uint64_t cap;
    uint32_t *nev;             // per elected variable: its event count
};

__device__ __forceinline__ void plog(const ProofLog &L, uint32_t q, uint32_t &seq, uint32_t tag, uint64_t payload)
{
    if (!L.ev) return;
    const unsigned long long i = atomicAdd(L.n, 1ull);
    if (i < L.cap) {
        PEv e;
        e.q = q;
        e.seq = seq;
        e.tag = tag;
        e.pad = 0;
        e.payload = payload;
        L.ev[i] = e;
    }
    seq++;
}

// an event whose place in its variable's loop is known (seq), logged later
__device__ __forceinline__ void plog_at(const ProofLog &L, uint32_t q, uint32_t seq, uint32_t tag, uint64_t payload)
{
    if (!L.ev) return;
    const unsigned long long i = atomicAdd(L.n, 1ull);
    if (i < L.cap) {
        PEv e;
        e.q = q;
        e.seq = seq;
        e.tag = tag;
        e.pad = 0;
        e.payload = payload;
        L.ev[i] = e;
    }
}

__device__ __forceinline__ void plog_end(const ProofLog &L, uint32_t q, uint32_t seq)
{
    if (L.ev) L.nev[q] = seq;
}

}  // namespace pfg
