#pragma once

// =============================================================================
// naga/cospan/cospan.hpp — the whole CoSpan API of this backend (header-only; compiled by hipcc for gfx950)
// =============================================================================

#include "naga/cospan/component.hpp"
#include "naga/cospan/device.hpp"
#include "naga/cospan/dtype.hpp"
#include "naga/cospan/fused_component.hpp"
#include "naga/cospan/kernel.hpp"
#include "naga/cospan/model.hpp"
#include "naga/cospan/system.hpp"
#include "naga/cospan/variable.hpp"
#include "naga/cospan/vocab.hpp"
#include "naga/cospan/wrappers.hpp"
