"""TEST INFRASTRUCTURE ONLY: CPU oracle of the gfx950 kernels (the checker)."""
