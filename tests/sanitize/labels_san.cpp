// Host sanitizer run of the library's device-free entry points
// (lbm355_case_labels, lbm355_case_footprint, descriptor validation), built
// from the library's host half with -fsanitize=address,undefined: the label
// setup, open-boundary extrusion, brick index with its origin shift and the
// boundary partner tables on the channel, junctions, sizes that are no
// multiple of the brick, the smallest grids and random wall masks, plus the
// descriptor errors.  Exit status 0 and "labels_san: ok" when no check or
// sanitizer fired.
#include <cstdio>
#include <vector>

#include "cases.h"

static int failures = 0;

static void check(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "labels_san: %s failed (%s)\n", what, lbm355_last_error());
        ++failures;
    }
}

static void labels(const Mask& m, const lbm355_case_desc& d, int want)
{
    std::vector<int32_t> lab(size_t(m.nx) * m.ny * m.nz, -1);
    lbm355_case_info info;
    const int rc = lbm355_case_labels(&d, lab.data(), &info);
    check(rc == want, "case_labels return code");
    if (rc == LBM355_OK) {
        int64_t walls = 0;
        for (int32_t v : lab) {
            check(v >= 0 && v <= 8, "label range");
            walls += v == 1;
        }
        check(walls == info.wall, "wall count");
        int64_t bytes = 0;
        check(lbm355_case_footprint(&d, &bytes) == LBM355_OK && bytes == info.device_bytes, "footprint");
        check(lbm355_case_labels(&d, nullptr, nullptr) == LBM355_OK, "case_labels without outputs");
    }
}

int main()
{
    const Mask ch = channel(64, 32, 32, 12.0);
    lbm355_case_desc d = describe(ch, LBM355_CUM, LBM355_F32);
    channel_faces(d);
    labels(ch, d, LBM355_OK);
    for (int st : {LBM355_F64, LBM355_F32, LBM355_F16_OFFSET}) {
        const Mask j = junction(64, 48, 32, 5.0, 4.0, 3.5, 1.5);
        labels(j, describe(j, LBM355_CUM, st), LBM355_OK);
    }
    labels(junction(37, 29, 11, 4.0, 3.0, 3.0, 1.0), describe(junction(37, 29, 11, 4.0, 3.0, 3.0, 1.0), LBM355_KBC,
           LBM355_F32), LBM355_OK);
    std::mt19937 rng(7);
    const int sizes[][3] = {{4, 4, 4}, {4, 5, 7}, {9, 9, 4}, {17, 10, 6}, {33, 24, 9}, {8, 64, 5}, {70, 4, 4}};
    for (auto& s : sizes)
        for (double p : {0.2, 0.6, 1.0}) {
            const Mask m = random_mask(rng, s[0], s[1], s[2], p);
            lbm355_case_desc r = describe(m, LBM355_BGK, LBM355_F32);
            std::vector<int32_t> lab(m.v.size());
            lbm355_case_info info;
            const int rc = lbm355_case_labels(&r, lab.data(), &info);
            check(rc == LBM355_OK || rc == LBM355_ETERMINATE, "random mask labels");
            channel_faces(r);
            const int rc2 = lbm355_case_labels(&r, lab.data(), &info);
            check(rc2 == LBM355_OK || rc2 == LBM355_ETERMINATE, "random mask labels (channel faces)");
        }
    // descriptor errors
    lbm355_case_desc bad = describe(ch, LBM355_CUM, LBM355_F32);
    bad.nx = 3;
    labels(ch, bad, LBM355_EINVAL);
    bad = describe(ch, LBM355_CUM, 9);
    labels(ch, bad, LBM355_EINVAL);
    bad = describe(ch, LBM355_CUM, LBM355_F32);
    bad.geometry = nullptr;
    labels(ch, bad, LBM355_EINVAL);
    bad = describe(ch, LBM355_CUM, LBM355_F32);
    bad.face_bc[2] = LBM355_FACE_INFLOW;
    labels(ch, bad, LBM355_EINVAL);
    const Mask closed = random_mask(rng, 12, 12, 6, 0.0);
    labels(closed, describe(closed, LBM355_CUM, LBM355_F32), LBM355_ETERMINATE);
    if (failures) return 1;
    std::printf("labels_san: ok\n");
    return 0;
}
