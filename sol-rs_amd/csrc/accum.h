// accum.h -- order-free pixel sums for per-sample random streams.
//
// With sample_streams the samples of a pixel are traced in any order (many
// per pass, or on several server slots at once), so their radiance is summed
// in fixed point: each sample's value times 2^20, rounded to nearest (ties to
// even), as a 64-bit integer.  Integer addition does not depend on the order,
// so every way of splitting the samples gives the same bits.  A NaN, an
// infinity or a value of 2^19 or more marks the sum bad (so that 2^24 samples
// cannot overflow), and a bad pixel keeps its previous colour, as a NaN or
// infinite float sum does in pathtrace.rgen.
// The mean is (float)((double)sum * 2^-20) times 1/spp in float.
#pragma once

#include <stdint.h>

#include "devmath.h"

namespace solhip {

struct FixSum {         // 32 B
    long long c[3];
    uint32_t bad;       // a sample could not be represented
    uint32_t count;     // server: slots of the pixel that have finished
};

SOLHIP_HD void fix_clear(FixSum &s)
{
    s.c[0] = s.c[1] = s.c[2] = 0;
    s.bad = 0;
    s.count = 0;
}

SOLHIP_HD void fix_add1(FixSum &s, int k, float v)
{
    if (!(fabsf(v) < 0x1p19f)) {
        s.bad = 1;
        return;
    }
    s.c[k] += (long long)rintf(v * 0x1p20f);
}

SOLHIP_HD void fix_add(FixSum &s, float r, float g, float b)
{
    fix_add1(s, 0, r);
    fix_add1(s, 1, g);
    fix_add1(s, 2, b);
}

SOLHIP_HD void fix_merge(FixSum &s, const FixSum &o)
{
    s.c[0] += o.c[0];
    s.c[1] += o.c[1];
    s.c[2] += o.c[2];
    s.bad |= o.bad;
}

SOLHIP_HD float fix_value(long long c) { return (float)((double)c * 0x1p-20); }

}  // namespace solhip
