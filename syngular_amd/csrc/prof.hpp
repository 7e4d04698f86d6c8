// Event-based per-kernel timing used by syn_prof_begin / syn_prof_end.
#pragma once
#include <hip/hip_runtime.h>

namespace syn {

// kernel tags (indices of the arrays filled by syn_prof_end)
enum ProfTag {
    kTagGemmSingle64 = 0,
    kTagGemmSingle128 = 1,
    kTagGemmGrouped64 = 2,
    kTagGemmGrouped128 = 3,
    kTagBJacobiStep = 4,
    kTagCholDiag = 5,
    kTagQr = 6,
    kTagJacobi = 7,
    kTagApply = 8,
    kTagGramStep = 9,      // the fused Jacobi step (tjf_step_kernel)
    kTagGramBytes = 10,    // bytes its update work-groups move (work only, no launches)
    kTagGemmMixed64 = 11,
    kTagGemmMixed128 = 12,
    kTagGemmSplitK128 = 13,
    kTagGemmSplitK64 = 14,
    kTagGemmPair64 = 15,
    kTagGemmPair128 = 16,
    kTagCholNode = 17,
    kTagGemmBf1664 = 18,
    kTagGemmBf16128 = 19,
};

bool prof_on();
// Profiling is on and the tag is selected for event timing.
bool prof_wants(int tag);

// Adds work (flops or bytes) to a tag's total without events: for kernels
// whose work is only known on the device and summed up after the call.
void prof_add_work(int tag, double work);

// Records an event before (constructor) and after (destructor) the launch
// issued in its scope.  flops < 0 takes the value given by the last
// syn_prof_note_flops call.
class ProfScope {
   public:
    ProfScope(int tag, double flops, hipStream_t s);
    ~ProfScope();

   private:
    int tag_;
    double flops_;
    hipStream_t s_;
    hipEvent_t a_;
};

}  // namespace syn
