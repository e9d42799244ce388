// Line prefixes of the reference's layout files (include/hibf/layout/prefixes.hpp
// in the reference): the format is shared, so the strings are the same.
#pragma once

#include <string_view>

namespace seqan::hibf::prefix
{

constexpr std::string_view meta_header{"@"};
constexpr std::string_view meta_hibf_config_start{"@HIBF_CONFIG"};
constexpr std::string_view meta_hibf_config_end{"@HIBF_CONFIG_END"};
constexpr std::string_view layout_header{"#"};
constexpr std::string_view layout_top_level{"TOP_LEVEL_IBF"};
constexpr std::string_view layout_lower_level{"LOWER_LEVEL_IBF"};
constexpr std::string_view layout_fullest_technical_bin_idx{"fullest_technical_bin_idx:"};
constexpr std::string_view layout_first_header_line{"#TOP_LEVEL_IBF"};
constexpr std::string_view layout_column_names{"#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS"};

} // namespace seqan::hibf::prefix
