// Stream ingestion and the reference pipe's output formats.  See pipe.hip.
#pragma once

#include <stdint.h>

#include <string>

namespace ypdq {

// One output record format of the reference's `pipe -f` (src/main.rs:1106-1133,
// 1264-1290): raw / q+raw / hex / HEX / q+hex / q+HEX / bin / q+bin /
// diagnostic, optionally followed by +nl / +lf / +crlf / +cr / +nul.
struct OutputFormat {
  enum Type { kRaw, kRawQuality, kHex, kHexQuality, kBin, kBinQuality, kDiagnostic };
  Type type = kBin;
  bool upper = false;     // HEX
  std::string separator;  // written after every record
};

bool parse_output_format(const char* spec, OutputFormat* f);

// Records of n frames: hashes n x 8 x 32, quality and threshold n x 8 (the
// fused kernel's outputs, orientation 0 = the original), dihedral_pdqf
// n x 8 x 256 (the diagnostic format only, else may be null).  One record per
// frame, or eight (original, then visit_dihedrals' order) with dihedrals.
void format_frames(const OutputFormat& f, uint32_t n_frames, bool dihedrals, const uint8_t* hashes,
                   const float* quality, const float* threshold, const float* dihedral_pdqf, std::string& out);
// An upper bound of the bytes format_frames writes.
uint64_t format_bound(const OutputFormat& f, uint32_t n_frames, bool dihedrals);

constexpr int kPipeEio = -7;  // a read or write failed
constexpr int kPipeEof = -8;  // the input ended inside a frame

struct PipeOptions {
  int input_fd = 0;
  int output_fd = 1;
  int input_kind = 1;  // kInputU8 (luma8), kInputRgb8, kInputRgba8
  OutputFormat format;
  bool dihedrals = false;
  uint32_t dihedral_mode = 1;  // the x86 reference binary's AVX2 column extent
  uint32_t batch = 64;         // frames per GPU batch
  bool stats = false;          // a progress line on stderr about once a second
};

// Hashes every whole frame of the input, writing its records in order.
// Returns 0 at a clean end of input (a frame boundary), kPipeEof after
// writing the whole frames when the input ends inside one, kPipeEio on a
// read/write error, -1 for bad options, -2 for a HIP error (message in
// *error).  *frames_out: frames written.
int run_pipe(const PipeOptions& o, uint64_t* frames_out, std::string* error);

// Frees the buffer sets that finished runs keep for the next run of the same
// shape; returns how many were freed.
size_t trim_pipe_cache();

}  // namespace ypdq
