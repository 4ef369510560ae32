// Which group of systems a slot of the cooperative schedule solves.  Plain C++
// shared by the kernels and the host checks.
#pragma once

#if defined(__HIPCC__)
#define TRI_COOP_HD __host__ __device__
#else
#define TRI_COOP_HD
#endif

namespace tri {

// Group solved by slot s (tasks s*K .. s*K + K-1) of G groups.  A round of the
// persistent grid runs W = P/K slots at once.  With W > 1, the W slots of a round
// take groups R = ceil(G/W) apart: slot s = r*W + c is entry (row r, column c) of a
// column-major table of W columns, the first L = G - (R-1)*W of which hold R
// entries and the rest R - 1.  A round's loads then cover the whole width of an
// interleaved row instead of a strip of W*S neighbouring systems.  Measured with
// the kernels' own load pattern at FP32 4096 x 4096 (128-byte row pieces, 32
// groups per round): 3.2 TB/s with neighbouring groups, 3.9 TB/s spread.  The map
// is a bijection on [0, G) for every W; which slots share a round is unchanged,
// so every wait of the schedule still ends.
inline TRI_COOP_HD int coop_group(int s, int G, int W) {
    if (W <= 1 || G <= W) return s;
    const int R = (G + W - 1) / W;
    const int r = s / W, c = s - r * W;
    const int L = G - (R - 1) * W;
    return c < L ? c * R + r : L * R + (c - L) * (R - 1) + r;
}

}  // namespace tri
