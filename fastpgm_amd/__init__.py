"""MI355X-native PC-stable CI stage (G^2 skeleton search)."""
