// Specialised sub-FFT passes with compile-time shapes (auto_shape) for the further lengths that conversions between
// the common audio rates (8 kHz to 192 kHz, 12, 24 and 64 kHz included, both modes) need for inputs of 0.5 s to 600 s
// (short jobs of a batch among them) and the other tables lack; part 4.
#include "fixed_impl.hpp"

namespace kbsrc::gpu {

namespace {

const FixedEntry kAuto[] = {
    KB_AUTO(POL_DDS, 49, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 49, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 50, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 50, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 56, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 56, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 70, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 70, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 70, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 72, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 75, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 80, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 81, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 125, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 144, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 162, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 175, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 180, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 189, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 196, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 243, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 243, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 243, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 256, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 270, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 288, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 324, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 343, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 350, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 350, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 360, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 378, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 378, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 384, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_DDS, 384, LD_CPLX, ST_OUT),
    KB_AUTO(POL_DDS, 384, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_DDS, 405, LD_CPLX, ST_OUT),
    KB_AUTO(POL_F64, 50, LD_CPLX, ST_CPLX),
    KB_AUTO(POL_F64, 50, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_F64, 75, LD_REAL_F64, ST_CPLX),
    KB_AUTO(POL_F64, 192, LD_CPLX, ST_OUT),
};

}  // namespace

const FixedEntry* fixed_table_auto4(int* count) {
    *count = (int)(sizeof(kAuto) / sizeof(kAuto[0]));
    return kAuto;
}

}  // namespace kbsrc::gpu
