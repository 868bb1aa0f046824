// Sensor-side inputs in the reference's formats: the calibration file
// (calib.xml) -> calibration block, PNM frame files, and the Kinect
// disparity -> depth conversion (a HIP kernel).
#include <hip/hip_runtime.h>

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "rt_host.hpp"
#include "rt_sensor.h"

namespace rt {

// ------------------------------------------------------------------ calib.xml
// One element of the file: its size attribute ("cols rows") and its text.
struct XmlElem {
    bool found = false;
    int sx = 0, sy = 0;
    std::string text;
};

// First element named `name` (exact tag name) anywhere in the document; the
// reference takes the root's first child of that name (calib.cpp:571-692),
// and the format has no nesting below the root.
static XmlElem find_elem(const std::string& s, const char* name) {
    XmlElem e;
    const std::string open = std::string("<") + name;
    for (size_t pos = s.find(open); pos != std::string::npos; pos = s.find(open, pos + 1)) {
        const size_t after = pos + open.size();
        if (after >= s.size()) break;
        const char c = s[after];
        if (!(c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '>' || c == '/')) continue;
        const size_t gt = s.find('>', pos);
        if (gt == std::string::npos) break;
        const std::string tag = s.substr(pos, gt - pos);
        const size_t sa = tag.find("size=\"");
        if (sa != std::string::npos) sscanf(tag.c_str() + sa + 6, "%d %d", &e.sx, &e.sy);
        e.found = true;
        if (!tag.empty() && tag.back() == '/') return e;  // <name ... />
        const size_t close = s.find(std::string("</") + name, gt);
        e.text = s.substr(gt + 1, (close == std::string::npos ? s.size() : close) - gt - 1);
        return e;
    }
    return e;
}

// n numbers of an element's text, separated by white space (parseDoubles,
// calib.cpp:25-38, reads them with atof one after the other)
static bool parse_doubles(const std::string& text, int n, double* out) {
    const char* p = text.c_str();
    for (int i = 0; i < n; i++) {
        char* end = nullptr;
        out[i] = strtod(p, &end);
        if (end == p) return false;
        p = end;
    }
    return true;
}

// transpose (calib.cpp:66-71): the file stores the 3x3 matrices column by column
static void transpose3(const double* P, double* M) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M[r * 3 + c] = P[c * 3 + r];
}

// Rotation matrix of a rotation vector (the Rodrigues formula the reference
// obtains from cv::Rodrigues, calib.cpp:649-656): R = cos t I + (1 - cos t) k k^T
// + sin t [k]x with t = |om|, k = om / t; the identity for t = 0.
static void rodrigues(const double* om, double* R) {
    const double t = sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
    if (t < 2.220446049250313e-16) {
        for (int i = 0; i < 9; i++) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    const double kx = om[0] / t, ky = om[1] / t, kz = om[2] / t;
    const double c = cos(t), s = sin(t), C = 1.0 - c;
    R[0] = c + kx * kx * C;      R[1] = kx * ky * C - kz * s; R[2] = kx * kz * C + ky * s;
    R[3] = ky * kx * C + kz * s; R[4] = c + ky * ky * C;      R[5] = ky * kz * C - kx * s;
    R[6] = kz * kx * C - ky * s; R[7] = kz * ky * C + kx * s; R[8] = c + kz * kz * C;
}

// det3x3 / inverse3x3 in float (basic_math.cpp:507-514, 562-582): the
// identity when |det| < 1e-5
static void inverse3x3f(const float* m, float* r) {
    const float det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) +
                      m[2] * (m[3] * m[7] - m[6] * m[4]);
    if (fabsf(det) < 1e-5f) {
        for (int i = 0; i < 9; i++) r[i] = (i % 4 == 0) ? 1.0f : 0.0f;
        return;
    }
    float t[9];
    t[0] = (m[4] * m[8] - m[5] * m[7]) / det;
    t[1] = -(m[1] * m[8] - m[7] * m[2]) / det;
    t[2] = (m[1] * m[5] - m[4] * m[2]) / det;
    t[3] = -(m[3] * m[8] - m[5] * m[6]) / det;
    t[4] = (m[0] * m[8] - m[6] * m[2]) / det;
    t[5] = -(m[0] * m[5] - m[3] * m[2]) / det;
    t[6] = (m[3] * m[7] - m[6] * m[4]) / det;
    t[7] = -(m[0] * m[7] - m[6] * m[1]) / det;
    t[8] = (m[0] * m[4] - m[1] * m[3]) / det;
    memcpy(r, t, sizeof(t));
}

// setupCalibDataBuffer (calib.cpp:73-93): intrinsics rows 0-1 scaled by the
// width (the file normalises both by it), row 2 = (0, 0, 1)
static void intrinsics_block(const double* K, int width, float* out, float* inv) {
    for (int i = 0; i < 9; i++) out[i] = (i % 4 == 0) ? 1.0f : 0.0f;
    for (int i = 0; i < 6; i++) out[i] = (float)(width * K[i]);
    inverse3x3f(out, inv);
}

static int parse_calib(const std::string& s, int width, int height, float* calib, int* xy_offset) {
    if (width <= 0 || height <= 0 || !calib) return set_error(RT_ERR_ARG, "calibration: bad argument");
    double tmp[16], KL[9], KR[9], kcR[8] = {0}, B = 0.0, b = 0.0, om[3], T[3], R[9];
    XmlElem e = find_elem(s, "KL");
    if (!e.found || e.sx * e.sy != 9 || !parse_doubles(e.text, 9, tmp)) return set_error(RT_ERR_ARG, "calibration: KL not found");
    transpose3(tmp, KL);
    e = find_elem(s, "KR");
    if (!e.found || e.sx * e.sy != 9 || !parse_doubles(e.text, 9, tmp)) return set_error(RT_ERR_ARG, "calibration: KR not found");
    transpose3(tmp, KR);
    e = find_elem(s, "kc_right");
    if (e.found) {
        const int n = e.sx * e.sy;
        if (n < 1 || n > 8 || !parse_doubles(e.text, n, kcR)) return set_error(RT_ERR_ARG, "calibration: bad kc_right");
    }
    e = find_elem(s, "B");
    if (e.found && !parse_doubles(e.text, 1, &B)) return set_error(RT_ERR_ARG, "calibration: bad B");
    e = find_elem(s, "b");
    if (e.found && !parse_doubles(e.text, 1, &b)) return set_error(RT_ERR_ARG, "calibration: bad b");
    e = find_elem(s, "om");
    if (!e.found || e.sx * e.sy != 3 || !parse_doubles(e.text, 3, om)) return set_error(RT_ERR_ARG, "calibration: om not found");
    e = find_elem(s, "T");
    if (!e.found || e.sx * e.sy != 3 || !parse_doubles(e.text, 3, T)) return set_error(RT_ERR_ARG, "calibration: T not found");
    float zpoly[16] = {0};
    int nz = 0;
    e = find_elem(s, "zpoly");
    if (e.found) {
        nz = e.sx * e.sy;
        if (nz < 0 || nz > 16 || !parse_doubles(e.text, nz, tmp)) return set_error(RT_ERR_ARG, "calibration: bad zpoly");
        for (int i = 0; i < nz; i++) zpoly[i] = (float)tmp[i];
    }
    // TLR = [R(om) | T], TRL its inverse (calib.cpp:644-674)
    double TLR[16], TRL[16];
    rodrigues(om, R);
    for (int i = 0; i < 16; i++) TLR[i] = (i % 5 == 0) ? 1.0 : 0.0;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) TLR[r * 4 + c] = R[r * 3 + c];
    TLR[3] = T[0]; TLR[7] = T[1]; TLR[11] = T[2];
    invert_rt4(TLR, TRL);
    // disparity constants (calib.cpp:699-701): z = 1 / (c0 + c1 d)
    const double scale = 1.0f / (8 * b * fabs(KL[0]) * 640.0);
    const double c1 = -scale, c0 = B * 640.0f * scale;

    memset(calib, 0, sizeof(float) * RT_CALIB_SIZE);
    intrinsics_block(KR, width, calib + RT_KR_OFFSET, calib + RT_IKR_OFFSET);
    for (int i = 0; i < 8; i++) calib[RT_KCR_OFFSET + i] = (float)kcR[i];
    for (int i = 0; i < 16; i++) {
        calib[RT_TLR_OFFSET + i] = (float)TLR[i];
        calib[RT_TRL_OFFSET + i] = (float)TRL[i];
    }
    intrinsics_block(KL, width, calib + RT_KL_OFFSET, calib + RT_IKL_OFFSET);
    calib[RT_C0_OFFSET] = (float)c0;
    calib[RT_C1_OFFSET] = (float)c1;
    calib[RT_MIND_OFFSET] = 300.0f;
    calib[RT_MAXD_OFFSET] = 7000.0f;
    for (int i = 0; i < nz; i++) calib[80 + i] = zpoly[i];
    if (xy_offset) *xy_offset = 1;
    return RT_OK;
}

static bool read_file(const char* path, std::vector<char>& buf) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n < 0) { fclose(f); return false; }
    buf.resize((size_t)n);
    size_t got = n ? fread(buf.data(), 1, (size_t)n, f) : 0;
    fclose(f);
    return got == (size_t)n;
}

// ------------------------------------------------------------------ PNM
struct PnmHeader { int w = 0, h = 0, ch = 0, maxval = 0; size_t offset = 0; };

static bool pnm_header(const std::vector<char>& b, PnmHeader& hd) {
    if (b.size() < 3 || b[0] != 'P' || (b[1] != '5' && b[1] != '6')) return false;
    hd.ch = b[1] == '5' ? 1 : 3;
    size_t i = 2;
    int vals[3];
    for (int k = 0; k < 3; k++) {
        // white space and comments before each number
        for (;;) {
            while (i < b.size() && isspace((unsigned char)b[i])) i++;
            if (i < b.size() && b[i] == '#') { while (i < b.size() && b[i] != '\n') i++; continue; }
            break;
        }
        if (i >= b.size() || !isdigit((unsigned char)b[i])) return false;
        long v = 0;
        while (i < b.size() && isdigit((unsigned char)b[i])) { v = v * 10 + (b[i] - '0'); i++; if (v > 1 << 20) return false; }
        vals[k] = (int)v;
    }
    if (i >= b.size() || !isspace((unsigned char)b[i])) return false;
    i++;  // one white space character before the samples
    hd.w = vals[0]; hd.h = vals[1]; hd.maxval = vals[2]; hd.offset = i;
    return hd.w > 0 && hd.h > 0 && hd.maxval > 0 && hd.maxval < 65536;
}

// ------------------------------------------------------------------ disparity -> depth
// d2ZLow / d2Z / d2ZLowGPU (multicore.cpp:189-379), d2ZLowKernel (d2z.cu:117-160)
__global__ __launch_bounds__(256) void k_disparity_to_depth(const uint16_t* __restrict__ disp, int w, int h, int ow,
                                                            int oh, int half, float c0, float c1, float mn, float mx,
                                                            int ox, int oy, int mode, float* __restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= ow || y >= oh) return;
    float v = 0.0f;
    if (y < 4 * (oh / 4)) {  // the reference's four row blocks of oh/4 rows
        const int sx = (half ? 2 * x : x) + ox, sy = (half ? 2 * y : y) + oy;
        if (sx >= 0 && sy >= 0 && sx < w - 1 && sy < h - 1) {
            const uint16_t* p = disp + (size_t)sx + (size_t)sy * w;
            float d = (float)p[0];
            bool ok = d < 2047.0f;
            if (half) {
                const float d2 = (float)p[1], d3 = (float)p[w + 1], d4 = (float)p[w];
                ok = ok && d2 < 2047.0f && d3 < 2047.0f && d4 < 2047.0f;
                if (d2 < d) d = d2;
                if (d3 < d) d = d3;
                if (d4 < d) d = d4;
            }
            if (ok) {
                float z = fabsf(1.0f / (c0 + c1 * d));
                if (z > mx || z < mn) z = 0.0f;
                v = mode ? (z - mn) / (mx - mn) : z;
            }
        }
    }
    out[(size_t)x + (size_t)y * ow] = v;
}

__global__ __launch_bounds__(256) void k_depth_mm_to_u16(const float* __restrict__ z, size_t n, uint16_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float v = z[i];
    out[i] = (v > 0.0f && v < 65535.5f) ? (uint16_t)(v + 0.5f) : (uint16_t)0;
}

}  // namespace rt

using namespace rt;

extern "C" int rt_calib_parse_xml(const char* text, size_t len, int width, int height, float* calib, int* xy_offset) {
    if (!text) return set_error(RT_ERR_ARG, "calibration: null text");
    return parse_calib(std::string(text, len), width, height, calib, xy_offset);
}

extern "C" int rt_calib_load_xml(const char* path, int width, int height, float* calib, int* xy_offset) {
    std::vector<char> buf;
    if (!path || !read_file(path, buf)) return set_error(RT_ERR_ARG, "calibration: cannot read the file");
    return parse_calib(std::string(buf.begin(), buf.end()), width, height, calib, xy_offset);
}

extern "C" int rt_pnm_info(const char* path, int* width, int* height, int* channels, int* maxval) {
    std::vector<char> buf;
    PnmHeader hd;
    if (!path || !read_file(path, buf)) return set_error(RT_ERR_ARG, "pnm: cannot read the file");
    if (!pnm_header(buf, hd)) return set_error(RT_ERR_ARG, "pnm: not a binary P5/P6 file");
    if (width) *width = hd.w;
    if (height) *height = hd.h;
    if (channels) *channels = hd.ch;
    if (maxval) *maxval = hd.maxval;
    return RT_OK;
}

extern "C" int rt_pnm_read(const char* path, void* data, size_t bytes) {
    std::vector<char> buf;
    PnmHeader hd;
    if (!path || !data || !read_file(path, buf)) return set_error(RT_ERR_ARG, "pnm: cannot read the file");
    if (!pnm_header(buf, hd)) return set_error(RT_ERR_ARG, "pnm: not a binary P5/P6 file");
    const size_t bps = hd.maxval > 255 ? 2 : 1;
    const size_t n = (size_t)hd.w * hd.h * hd.ch;
    if (bytes < n * bps) return set_error(RT_ERR_ARG, "pnm: buffer too small");
    if (buf.size() < hd.offset + n * bps) return set_error(RT_ERR_ARG, "pnm: file shorter than its header says");
    const unsigned char* src = (const unsigned char*)buf.data() + hd.offset;
    if (bps == 1) {
        memcpy(data, src, n);
    } else {
        uint16_t* dst = (uint16_t*)data;
        for (size_t i = 0; i < n; i++) dst[i] = (uint16_t)(src[2 * i] << 8 | src[2 * i + 1]);  // big-endian on disk
    }
    return RT_OK;
}

extern "C" int rt_disparity_to_depth(const uint16_t* disparity, int width, int height, const float* calib, int xy_offset,
                                     int half, int mode, float* depth, void* stream) {
    if (!disparity || !calib || !depth || width < 2 || height < 2 || (mode != 0 && mode != 1))
        return set_error(RT_ERR_ARG, "rt_disparity_to_depth: bad argument");
    const int ow = half ? width / 2 : width, oh = half ? height / 2 : height;
    const int ox = xy_offset ? -4 : 0, oy = xy_offset ? -3 : 0;
    dim3 grid((ow + 15) / 16, (oh + 15) / 16);
    k_disparity_to_depth<<<grid, 256, 0, (hipStream_t)stream>>>(disparity, width, height, ow, oh, half ? 1 : 0,
                                                                calib[RT_C0_OFFSET], calib[RT_C1_OFFSET],
                                                                calib[RT_MIND_OFFSET], calib[RT_MAXD_OFFSET], ox, oy,
                                                                mode, depth);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_disparity_to_depth: launch");
    return RT_OK;
}

extern "C" int rt_depth_mm_to_u16(const float* depth, size_t n, uint16_t* out, void* stream) {
    if (!depth || !out) return set_error(RT_ERR_ARG, "rt_depth_mm_to_u16: bad argument");
    if (n == 0) return RT_OK;
    k_depth_mm_to_u16<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(depth, n, out);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_depth_mm_to_u16: launch");
    return RT_OK;
}
