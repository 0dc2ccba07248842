// Subset pixel lists, host side: LocalMask::new (masks.rs:58-187) and
// mask_update (masks.rs:199-263).  Integer work, bit-exact with the
// reference: row-major scan (row outer, column inner), [x_off, y_off]
// pairs, circle test sqrt(x²+y²) <= r, semicircle keeps y_off >= 0.
#include <cmath>
#include <string>
#include <vector>

#include "../../include/geodic.h"

namespace geodic {
void set_error(const std::string& msg);
}

namespace {

bool keep_px(int shape, long x, long y, int size) {
    if (shape == DIC_SHAPE_SQUARE) return true;
    const double dist = std::sqrt((double)(x * x + y * y));
    bool k = dist <= (double)size;
    if (shape == DIC_SHAPE_SEMICIRCLE) k = k && y >= 0;
    return k;
}

}  // namespace

extern "C" int dic_mask_local(int shape, int size, double* coords_out, int32_t* subset_mask_out, int64_t* n_out) {
    if (shape < 0 || shape > 2) {
        geodic::set_error("unknown mask shape");
        return DIC_ERR_INVALID_INPUT;
    }
    if (size <= 0) {
        // masks.rs:70-72 (and 115-117, 155-157)
        geodic::set_error(shape == DIC_SHAPE_SQUARE ? "invalid input: length must be > 0"
                                                    : "invalid input: radius must be > 0");
        return DIC_ERR_INVALID_INPUT;
    }
    if (size > 16383) {
        geodic::set_error("invalid input: mask size too large");
        return DIC_ERR_INVALID_INPUT;
    }
    const int side = 2 * size + 1;
    int64_t n = 0;
    for (int row = 0; row < side; ++row)
        for (int col = 0; col < side; ++col) {
            const long x = (long)col - size, y = (long)row - size;
            const bool k = keep_px(shape, x, y, size);
            if (subset_mask_out) subset_mask_out[(int64_t)row * side + col] = k ? 1 : 0;
            if (k) {
                if (coords_out) {
                    coords_out[2 * n] = (double)x;
                    coords_out[2 * n + 1] = (double)y;
                }
                ++n;
            }
        }
    if (n_out) *n_out = n;
    return DIC_OK;
}

extern "C" int dic_mask_update(int shape, int size, double cx, double cy, const uint8_t* mask, int64_t mh, int64_t mw,
                               double* coords_out, int64_t* n_out) {
    const int side = 2 * size + 1;
    std::vector<int32_t> sm((size_t)side * side);
    int st = dic_mask_local(shape, size, nullptr, sm.data(), nullptr);
    if (st) return st;
    // `as isize` truncates toward zero (masks.rs:204-205)
    const long center_x = (long)cx, center_y = (long)cy;
    long ly = center_y - size, uy = center_y + size + 1;
    long lx = center_x - size, ux = center_x + size + 1;
    long lyp = 0, uyp = 0, lxp = 0, uxp = 0;
    if (lx < 0) {
        lxp = -lx;
        lx = 0;
    }
    if (ly < 0) {
        lyp = -ly;
        ly = 0;
    }
    // an exclusive end at or past the edge also drops the last column/row
    // (masks.rs:225-232)
    if (ux >= mw) {
        uxp = ux + 1 - mw;
        ux = mw - 1;
    }
    if (uy >= mh) {
        uyp = uy + 1 - mh;
        uy = mh - 1;
    }
    const long trow_end = side - uyp, tcol_end = side - uxp;
    if (ly > uy || lx > ux || trow_end < lyp || tcol_end < lxp || uy - ly != trow_end - lyp ||
        ux - lx != tcol_end - lxp) {
        // the reference's ndarray slice/assign panics here
        geodic::set_error("invalid input: subset window lies outside the global mask");
        return DIC_ERR_INVALID_INPUT;
    }
    std::vector<uint8_t> local((size_t)side * side, 0);
    for (long r = 0; r < uy - ly; ++r)
        for (long c = 0; c < ux - lx; ++c)
            local[(size_t)(lyp + r) * side + (lxp + c)] = mask[(size_t)(ly + r) * mw + (lx + c)];
    int64_t n = 0;
    for (int row = 0; row < side; ++row)
        for (int col = 0; col < side; ++col)
            if (local[(size_t)row * side + col] != 0 && sm[(size_t)row * side + col] != 0) {
                if (coords_out) {
                    coords_out[2 * n] = (double)(col - size);
                    coords_out[2 * n + 1] = (double)(row - size);
                }
                ++n;
            }
    if (n_out) *n_out = n;
    return DIC_OK;
}
