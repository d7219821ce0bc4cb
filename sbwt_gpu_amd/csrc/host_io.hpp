// Host-side query text parsing shared by the C ABI's parsers and the
// pipelined text search.
#pragma once
#include <cstdint>

#include "../../include/sbwt_gpu_amd.h"

namespace sbwt {

int fail(int code, const char* fmt, ...);
void clear_error();

enum TextKind { kTextLines = 0, kTextFasta = 1, kTextFastq = 2 };

// The format of a text, from its first non-empty line: '>' FASTA, '@' FASTQ,
// otherwise one row per line.
TextKind text_kind(const char* text, uint64_t len);

// Rows of a text of the given kind (sbwt_parse_lines / sbwt_parse_records).
int parse_text(TextKind kind, const char* text, uint64_t len, uint8_t* chars, uint64_t* row_starts, uint64_t row_cap,
               uint64_t* rows_out);

// sbwt_parse_lines on host threads (parse_threads() of them for texts of 4
// MiB or more): the same rows and characters.
int parse_lines_parallel(const char* text, uint64_t len, uint8_t* chars, uint64_t* row_starts, uint64_t row_cap,
                         uint64_t* rows_out);

// Host threads the parsers use: the machine's, at most 16.
unsigned parse_threads();

// The first row boundary at or after `target` (a line start; for FASTA a
// '>' line; for FASTQ a record end, walked from the boundary `from`), or len.
uint64_t text_cut(TextKind kind, const char* text, uint64_t len, uint64_t from, uint64_t target, int* err);

}  // namespace sbwt
