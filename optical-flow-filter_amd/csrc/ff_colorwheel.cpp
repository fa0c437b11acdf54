// ff_colorwheel.cpp -- the colour wheel table of FlowToColor, embedded.
//
// data/colorwheel_rgba.bin holds the reference's wheel (256 x 256 RGBA,
// rows top to bottom; src/colorwheel.cpp:12, getColorWheelRGBA at
// src/colorwheel.cpp:21862-21873): the colours of FlowToColor are its
// texels, so they must be these exact bytes (data/colorwheel_extract.py
// explains where they come from). The assembler includes the file into the
// library's read-only data; the build passes its path as FF_WHEEL_BIN.
#include "ff_kernels.hpp"

#ifndef FF_WHEEL_BIN
#error "FF_WHEEL_BIN (path of data/colorwheel_rgba.bin) must be defined by the build"
#endif

asm(".section .rodata\n"
    ".balign 256\n"
    ".hidden ffamd_colorwheel_rgba\n"
    ".hidden ffamd_colorwheel_rgba_end\n"
    "ffamd_colorwheel_rgba:\n"
    ".incbin \"" FF_WHEEL_BIN "\"\n"
    "ffamd_colorwheel_rgba_end:\n"
    ".previous\n");

extern "C" const unsigned char ffamd_colorwheel_rgba[];
extern "C" const unsigned char ffamd_colorwheel_rgba_end[];

namespace ffamd {

const unsigned char* colorwheel_rgba(size_t* bytes) {
    if (bytes) *bytes = (size_t)(ffamd_colorwheel_rgba_end - ffamd_colorwheel_rgba);
    return ffamd_colorwheel_rgba;
}

}  // namespace ffamd
