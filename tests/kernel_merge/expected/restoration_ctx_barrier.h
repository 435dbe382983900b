#pragma once
struct Restoration_ctx {
  int target;
  int i;
};
