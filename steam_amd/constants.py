"""Model constants on the field-generation path (steam/constants.py:11-15)."""

hurst_horizontal = 0.45
haar_to_mhat = 0.25518
hurst_vertical_anisotropy = 5 / 9
