"""MI355X-native 4D modulated Henon-map tracker."""
