// resident.hip -- placeholder until the resident kernel lands.
#include "engine_internal.h"

namespace mte {

bool resident_fits(const mte_engine *, ResidentPlan *plan) {
    *plan = ResidentPlan{};
    return false;
}

hipError_t launch_resident(mte_engine *, int, bool) { return hipErrorNotSupported; }

}  // namespace mte
