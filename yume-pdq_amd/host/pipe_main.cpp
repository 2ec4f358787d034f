// yume-pdq-pipe: the reference's `pipe` subcommand (src/main.rs:140-252,
// 1091-1304) over the GPU stream-ingestion entry point.  Reads raw 512x512
// frames (luma8, rgb8 or rgba8) from a file or stdin and writes their hashes
// in the reference's output formats.
//
//   yume-pdq-pipe [-i PATH|-] [-o PATH|-] [-f FORMAT] [--input-format luma8|rgb8|rgba8]
//                 [--dihedrals] [--stats] [--batch N] [--exact-dihedrals]
//
// FORMAT: raw q+raw hex HEX q+hex q+HEX bin q+bin diagnostic, optionally
// followed by +nl +lf +crlf +cr +nul (default: bin).  --exact-dihedrals uses
// the full column negation of the scalar and AVX-512 kernels instead of the
// x86 AVX2 kernel's.  Exit status: 0 at a clean end of input, 5 on a read or
// write error or an input that ends inside a frame, 255 for bad arguments,
// 128 for a GPU error.
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <string>

#include "yume_pdq_amd.h"

static int usage(const char* msg) {
  fprintf(stderr, "%s\n", msg);
  fprintf(stderr,
          "usage: yume-pdq-pipe [-i PATH|-] [-o PATH|-] [-f FORMAT] [--input-format luma8|rgb8|rgba8]\n"
          "                     [--dihedrals] [--stats] [--batch N] [--exact-dihedrals]\n"
          "FORMAT: raw q+raw hex HEX q+hex q+HEX bin q+bin diagnostic [+nl|+lf|+crlf|+cr|+nul]\n");
  return 255;
}

int main(int argc, char** argv) {
  std::string in = "-", out = "-", fmt = "bin", input_format = "luma8";
  bool dihedrals = false, stats = false, exact = false;
  unsigned batch = 64;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto value = [&](const char* name) -> const char* {
      if (i + 1 >= argc) {
        usage((std::string("missing value for ") + name).c_str());
        exit(255);
      }
      return argv[++i];
    };
    if (a == "-i" || a == "--input") in = value("--input");
    else if (a == "-o" || a == "--output") out = value("--output");
    else if (a == "-f" || a == "--format") fmt = value("--format");
    else if (a == "--input-format") input_format = value("--input-format");
    else if (a == "--dihedrals") dihedrals = true;
    else if (a == "--stats") stats = true;
    else if (a == "--exact-dihedrals") exact = true;
    else if (a == "--batch") batch = (unsigned)strtoul(value("--batch"), nullptr, 10);
    else return usage(("unknown argument: " + a).c_str());
  }
  int kind;
  if (input_format == "luma8" || input_format == "LUMA8") kind = YUME_PDQ_AMD_INPUT_U8;
  else if (input_format == "rgb8" || input_format == "RGB8") kind = YUME_PDQ_AMD_INPUT_RGB8;
  else if (input_format == "rgba8" || input_format == "RGBA8") kind = YUME_PDQ_AMD_INPUT_RGBA8;
  else return usage(("invalid input format: '" + input_format + "'").c_str());
  if (yume_pdq_amd_format_bound(fmt.c_str(), 1, 0) == 0)
    return usage(("invalid output format: '" + fmt + "'").c_str());
  if (batch == 0 || batch > 4096) return usage("--batch must be 1..4096");

  const int in_fd = in == "-" ? 0 : open(in.c_str(), O_RDONLY);
  if (in_fd < 0) {
    perror(in.c_str());
    return 5;
  }
  const int out_fd = out == "-" ? 1 : open(out.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (out_fd < 0) {
    perror(out.c_str());
    return 5;
  }
  YumePdqAmdPipeConfig cfg;
  cfg.input_fd = in_fd;
  cfg.output_fd = out_fd;
  cfg.input_kind = kind;
  cfg.output_format = fmt.c_str();
  cfg.dihedrals = dihedrals;
  cfg.dihedral_mode = exact ? YUME_PDQ_AMD_DIHEDRAL_EXACT : YUME_PDQ_AMD_DIHEDRAL_AVX2;
  cfg.batch_frames = batch;
  cfg.stats = stats;
  uint64_t frames = 0;
  const int rc = yume_pdq_amd_pipe_run(&cfg, &frames);
  if (in_fd != 0) close(in_fd);
  if (out_fd != 1 && close(out_fd) != 0 && rc == YUME_PDQ_AMD_OK) {
    perror(out.c_str());
    return 5;
  }
  if (rc == YUME_PDQ_AMD_OK) return 0;
  fprintf(stderr, "yume-pdq-pipe: %s after %llu frames\n", yume_pdq_amd_strerror(rc), (unsigned long long)frames);
  if (rc == YUME_PDQ_AMD_EIO || rc == YUME_PDQ_AMD_EEOF) return 5;
  if (rc == YUME_PDQ_AMD_EINVAL) return 255;
  return 128;
}
