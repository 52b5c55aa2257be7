// Symbol visibility for the shared libraries: everything is hidden unless marked.
#pragma once
#define GAIL_AMD_EXPORT __attribute__((visibility("default")))
