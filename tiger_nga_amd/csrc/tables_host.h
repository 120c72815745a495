// Host-side construction of a mechanism's device table block.
#pragma once
#include <string>
#include <vector>

#include "../../include/tiger_chem.h"
#include "mech_tables.h"

namespace tgr {

// Fills `blob` with the table block of mechanism `t` (layout TabLayout for
// its compile-time shape) and `info` with its sizes.  Returns 0, or -1 with
// a message in *err when no kernel instance matches the mechanism.
int build_table_blob(const tgr_mech_tables* t, std::vector<char>& blob, tgr_mech_info* info, std::string* err);

}  // namespace tgr
