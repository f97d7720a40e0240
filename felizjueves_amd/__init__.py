"""MI355X-native two-electron Fock build (gfx950 HIP + ctypes C-ABI)."""
