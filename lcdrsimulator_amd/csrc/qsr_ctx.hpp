// qsr_ctx.hpp -- what other parts of the library may see of a solver context
// (defined in qsr_capi.cpp): its device field, one box's description and an
// explicit-Euler residual that leaves the field untouched.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "qsr_kernels.hpp"

namespace qsrc {

struct FieldView {
    int device = 0;
    hipStream_t st = nullptr;
    const float* cur = nullptr;  // the current field (five SoA planes of cs floats)
    float* other = nullptr;      // the other ping-pong buffer
    const float* rhs = nullptr;  // one explicit step of the current field, after euler_residual (else null)
    long long cs = 0, nvox = 0;
    qsrk::BoxDev box{};
    const qsrk::RowMask* rows = nullptr;  // step row masks (packed over all boxes)
    // the box's step row masks on the host, 8 int16 per row (i, j) at (i * Ny + j) * 8: the
    // interval [e0, e1] of k that evolves first (empty when e0 > e1)
    const int16_t* host_rows = nullptr;
    int nboxes = 0;
    bool uploaded = false;
    bool active = true;  // held boxes are not stepped (their step buffer is not written)
};

// Fails with QSR_E_ARG for a bad context or box index.
int field_view(void* ctx, int box, FieldView* out);
// Makes the other buffer current.  Only for single-box contexts, after every
// voxel a step can change has been written into it (the rest is identical in
// both buffers).
int commit_other(void* ctx);
// sqrt of the sum over the box's changeable interior voxels of |Q' - Q|^2,
// Q' = one explicit Euler step without noise of the current field (the
// semi-implicit right-hand side with no stabiliser); the field is unchanged.
// Q' stays in FieldView::rhs until the next step; count (may be null)
// receives the number of voxels summed over.
int euler_residual(void* ctx, int box, double* res, long long* count = nullptr);
// The two halves of euler_residual: one explicit step of every active box into
// FieldView::rhs, then the residual of one box from it.
int euler_rhs(void* ctx);
int box_residual(void* ctx, int box, double* res, long long* count = nullptr);
// box_residual of n boxes with one host synchronisation (counts may be null).
int box_residuals(void* ctx, const int* boxes, int n, double* res, long long* counts);
// Batched contexts: swap the ping-pong buffers of every box (the caller makes
// sure both hold what it wants for every box); copy the current field into
// the other buffer; copy one box's region of the other buffer into the current.
int swap_buffers(void* ctx);
int copy_current_to_other(void* ctx);
int copy_box_other_to_current(void* ctx, int box);
// Sets the calling thread's qsr_last_error() message and returns code.
int set_error(int code, const std::string& msg);

}  // namespace qsrc
