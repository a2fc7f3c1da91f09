// extern "C" access to the reference's own z_index translation unit
// (emerald/z_index/z_index.{h,cpp}), compiled from the reference tree by
// oracle/Makefile into oracle/_ref/. Test infrastructure only: it lets the
// tests check the restated Morton code against the reference's compiled code.
#include <emerald/z_index/z_index.h>

#include <cstdint>

extern "C" {

// z_index.h:8 -> z_index.cpp:5-10 (LUT-based Z_index_2::encode)
uint64_t ref_z_index_2(uint32_t x, uint32_t y) { return emerald::z_index::z_index(x, y); }

// z_index.h:29 Z_index_2::single_axis_index_from_lut
uint64_t ref_single_axis_index_2(uint32_t i) {
    static const emerald::z_index::Z_index_2 z;
    return z.single_axis_index_from_lut(i);
}

// z_index.h:11-16 Single_axis_index<2, 31>::encode (the template restatement
// the reference's own tests compare the LUT against)
uint64_t ref_single_axis_template_2(uint32_t i) {
    return emerald::z_index::Single_axis_index<2, 31>::encode(i);
}

}  // extern "C"
