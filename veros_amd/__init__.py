"""MI355X-native mixing hot path for veros (placeholder during bring-up)."""
