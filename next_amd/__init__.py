"""MI355X-native Barnes-Hut gravity + KDK time-step engine (NEXT drop-in)."""
