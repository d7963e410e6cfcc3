// AC tokenization in two halves, for hosts that sequence several stages
// before their one wait: the launches, then the check of the status words
// that the launches leave in GJXL_AMD_TOKENS_STATUS.
#pragma once

#include "common.h"

namespace gjxl_amd {

// Enqueues the four tokenization passes on the tokenizer's device stream
// (gjxl_amd_tokenize without its wait).
GJXLResult TokenizeEnqueue(GJXLAmdTokenizer* t, const int32_t* quantized,
                           const uint8_t* cells, const int32_t* raw_quant,
                           bool populations);

// Errors and the token count from the four status words of a finished
// tokenization, copied to the host.
GJXLResult TokenizeResult(const uint32_t status[4], uint64_t* token_count);

// The frame statistics of gjxl_amd_tokenizer_frame_statistics in two halves:
// the launches (their words land in GJXL_AMD_TOKENS_STATISTICS), then the
// check and unpacking of FrameStatisticsWords() words copied to the host.
GJXLResult FrameStatisticsEnqueue(GJXLAmdTokenizer* t, const int32_t* quantized,
                                  const uint8_t* cells, const int32_t* raw_quant);
GJXLResult FrameStatisticsResult(GJXLAmdTokenizer* t, const uint32_t* words,
                                 GJXLAmdFrameStatistics* out);
size_t FrameStatisticsWords();

// Re-targets the tokenizer to an extent whose buffers fit the ones it
// allocated; uploads that extent's sample decisions and waits.
GJXLResult ResizeTokenizer(GJXLAmdTokenizer* t, uint32_t width, uint32_t height);

}  // namespace gjxl_amd
