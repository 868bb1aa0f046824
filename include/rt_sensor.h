/*
 * rt_sensor.h -- the sensor-side inputs of the tracker in the reference's own
 * formats: its calibration file (calib.xml), its PNM frame files, and the
 * Kinect disparity -> depth conversion that turns raw disparity frames into
 * the depth maps the tracker consumes.
 *
 * Every entry point returns RT_OK (0) or a negative rt_status; rt_last_error()
 * gives the message.
 */
#ifndef RT_SENSOR_H
#define RT_SENSOR_H

#include <stddef.h>
#include <stdint.h>

#include "rt_tracker.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Calibration block (RT_CALIB_SIZE floats, the layout rt_create takes) from a
 * calibration file of the reference's XML format.
 * Replaces Calibration::init (tracker/calib/calib.cpp:521-716) followed by
 * Calibration::setupCalibDataBuffer(width, height) (calib.cpp:73-93): KL and KR
 * (stored column-major, normalised by the image width) scaled to `width`,
 * their float inverses, kc_right, TLR from the Rodrigues vector om and the
 * translation T, TRL its inverse, the disparity constants c0, c1 from B, b and
 * KL (calib.cpp:699-701), the depth window 300..7000 mm, and zpoly when the file
 * has one (entries 80.. of the block; zero otherwise). *xy_offset (may be
 * NULL) receives 1: files of this format put 2-D points on the IR image grid,
 * which needs the (-4, -3) disparity offset (calib.cpp:713).
 * A file without KL, KR, om or T is an error, as in the reference. */
int rt_calib_load_xml(const char* path, int width, int height, float* calib, int* xy_offset);
/* The same from the file's text in memory (len bytes). */
int rt_calib_parse_xml(const char* text, size_t len, int width, int height, float* calib, int* xy_offset);

/* Header of a binary PNM file (P5 gray or P6 RGB; maxval <= 255: 8-bit
 * samples, else 16-bit big-endian), the frame files the reference reads with
 * imread in FileSource::fetchRawImages (tracker/capture/fileSource.cpp:171-226):
 * bayer_rgbimageNNNN.ppm (8-bit Bayer mosaic) and rawdepthNNNN.ppm (11-bit
 * Kinect disparity, 2047 = no measurement). */
int rt_pnm_info(const char* path, int* width, int* height, int* channels, int* maxval);
/* The samples of a PNM file into data (bytes >= width*height*channels*(1 or
 * 2)); 16-bit samples are returned in host byte order. */
int rt_pnm_read(const char* path, void* data, size_t bytes);

/* Kinect disparity -> depth on the device (HIP kernel on `stream`, a
 * hipStream_t or NULL for the null stream; disparity and depth are device
 * pointers, calib a host pointer to the calibration block).
 * Replaces OMPFunctions::d2ZLow, d2Z and d2ZLowGPU
 * (tracker/multicore/multicore.cpp:189-379), the conversions ZConv::d2z runs
 * on every frame file (reconstruct/zconv.cpp:395-441), and the device twin
 * d2ZLowKernel (tracker/kernels/d2z.cu:117-160):
 *   half = 1: a (width/2) x (height/2) depth map; output pixel (x, y) takes the
 *             smallest of the four disparities at (2x + ox, 2y + oy) + {0,1}^2;
 *   half = 0: full resolution, the disparity at (x + ox, y + oy);
 * (ox, oy) = (-4, -3) when xy_offset, else (0, 0). A sample is valid when its
 * source lies inside the disparity image (a negative coordinate is outside, as
 * in the CPU conversion the file path uses) and every disparity is < 2047;
 * z = |1 / (c0 + c1 d)| mm, set to 0 outside the depth window [min, max].
 *   mode 0: depth in mm, 0 where invalid;
 *   mode 1: normalised (z - min) / (max - min), 0 where invalid (an
 *           out-of-window z gives -min / (max - min), as in the reference).
 * Rows at and below 4*floor(out_height/4) are 0 (the reference's four row
 * blocks never write them). depth holds out_width*out_height floats. */
int rt_disparity_to_depth(const uint16_t* disparity, int width, int height, const float* calib, int xy_offset,
                          int half, int mode, float* depth, void* stream);

/* Depth in mm (floats) -> the tracker's 16-bit depth input: rounded to the
 * nearest integer mm, 0 for z <= 0 or z >= 65535.5. This rounding is this
 * library's input format, not a step of the reference (which keeps the float
 * map); device pointers, n values, on `stream`. */
int rt_depth_mm_to_u16(const float* depth, size_t n, uint16_t* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RT_SENSOR_H */
