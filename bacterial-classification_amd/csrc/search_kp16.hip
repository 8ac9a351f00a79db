// search_kp16.hip -- search run functions with 16-entry candidate lists.
#include "search_impl.h"

namespace hvk {

HVK_SEARCH_RUNS(16, )
HVK_SEARCH_BOUNDED(16, )
}  // namespace hvk
