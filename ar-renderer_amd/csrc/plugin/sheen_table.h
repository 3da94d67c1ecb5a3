// Host-side sheen albedo table (sheen_table.cpp).
#pragma once

namespace pgpu {

// Writes the 16 x 32 table (alpha-major) of the Charlie sheen lobe's directional albedo.
void SheenAlbedoTable(float *out);

}  // namespace pgpu
