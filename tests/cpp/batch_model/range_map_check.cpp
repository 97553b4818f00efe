// range_of (gm_batch_plan.h) maps the workgroups of a batched launch onto its
// ranges: for every grid size G it must be a bijection of [0, G), and with
// G a multiple of the XCD count, consecutive ranges must mostly share an XCD
// (blocks of kXcdBlock) while every XCD gets an equal share.
//
//   range_map_check <Gmax>     exit status 0 when every check holds
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../../gpumatting_amd/csrc/gm_batch_plan.h"

using namespace gm::bplan;

int main(int argc, char** argv) {
    const int gmax = argc > 1 ? atoi(argv[1]) : 20000;
    long bad = 0;
    for (int G = 1; G <= gmax; ++G) {
        std::vector<int> seen(G, 0), per_xcd(kXcd, 0), xcd_of(G, -1);
        for (int b = 0; b < G; ++b) {
            const int r = range_of(b, G);
            if (r < 0 || r >= G || seen[r]++) {
                if (bad++ < 10) fprintf(stderr, "G %d: workgroup %d -> range %d (not a bijection)\n", G, b, r);
                break;
            }
            xcd_of[r] = b % kXcd;
            ++per_xcd[b % kXcd];
        }
        if (G % kXcd) continue;
        for (int x = 0; x < kXcd; ++x)
            if (per_xcd[x] != G / kXcd && bad++ < 10) fprintf(stderr, "G %d: XCD %d takes %d ranges\n", G, x, per_xcd[x]);
        // within a full block of kXcdBlock ranges, one XCD
        const int full = G / kXcd / kXcdBlock * kXcdBlock * kXcd;
        for (int r = 1; r < full; ++r)
            if (r % kXcdBlock && xcd_of[r] != xcd_of[r - 1] && bad++ < 10)
                fprintf(stderr, "G %d: ranges %d and %d of one block on XCDs %d, %d\n", G, r - 1, r, xcd_of[r - 1],
                        xcd_of[r]);
    }
    printf("range_of over G = 1..%d: %ld failed\n", gmax, bad);
    return bad ? 1 : 0;
}
