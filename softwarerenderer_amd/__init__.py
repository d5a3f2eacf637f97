"""MI355X (gfx950) implementation of trenki2/SoftwareRenderer's draw path."""
