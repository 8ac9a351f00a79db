// search_kp8.hip -- search run functions with 8-entry candidate lists.
#include "search_impl.h"

namespace hvk {

HVK_SEARCH_RUNS(8, )
HVK_SEARCH_BOUNDED(8, )
}  // namespace hvk
