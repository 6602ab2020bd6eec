// mc_tiff.h -- TIFF directories, strip/tile decoding and raw EER strips.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

struct TiffDir {
    int w = 0, h = 0;
    int bits = 0;        // BitsPerSample
    int fmt = 1;         // SampleFormat: 1 unsigned, 2 signed, 3 IEEE float
    int comp = 1;        // Compression
    int pred = 1;        // Predictor
    int spp = 1;         // SamplesPerPixel
    int rps = 0;         // RowsPerStrip (0: none given)
    int tw = 0, th = 0;  // tile size (0: strips)
    bool has_xres = false, has_yres = false;
    double xres = 0, yres = 0;
    std::vector<uint64_t> off, cnt;  // strip or tile offsets and byte counts
};

struct TiffFile {
    FILE* f = nullptr;
    uint64_t size = 0;  // file bytes
    bool be = false;    // big-endian file
    std::vector<TiffDir> dirs;
    std::string err;
    ~TiffFile() {
        if (f) fclose(f);
    }
};

// Open `path` and read every image file directory (classic and BigTIFF).
bool tiff_open(const char* path, TiffFile& t);
// MRC mode of a directory's samples as the reference maps them (0, 1, 6, 2),
// or -1.
int tiff_mode(const TiffDir& d);
// Decode directory `d` into `out` (w*h samples of tiff_mode's size, native
// byte order), rows flipped vertically as the reference stores them.
bool tiff_decode(TiffFile& t, const TiffDir& d, unsigned char* out);
// Raw (still encoded) bytes of every strip of `d`, back to back; `out` may be
// null to query the size.
bool tiff_raw(TiffFile& t, const TiffDir& d, unsigned char* out, uint64_t* bytes);
