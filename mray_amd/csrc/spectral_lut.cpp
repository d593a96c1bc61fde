// Spectral tables: built on first use (placeholder until the fitter lands).
#include "spectral.h"

namespace mray
{

int GetSpectralTables(SpectralTables& out)
{
    out.ready = false;
    SetError("spectral rendering: RGB-to-spectrum tables are not built yet");
    return MRAY_ERR_UNSUPPORTED;
}

} // namespace mray
