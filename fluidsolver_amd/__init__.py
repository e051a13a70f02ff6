"""MI355X-native SPH hot path: neighborhood search + IISPH/SESPH on gfx950."""
